/* pars_plan.h — the scan plan as a plain C value.
 *
 * This is the input contract of every pars backend: the output of the grammar
 * compiler (reference: include/pars/compile/plan.hpp:711-748) and the type the
 * reference backends take as a template argument (include/pars/plan.hpp:39-79).
 *
 * The layout is byte-for-byte the reference's `pars::scan_plan`: every member
 * is a one-byte type laid out in declaration order, so there is no padding and
 * sizeof(pars_plan) == 3177.  A caller holding a compiled `scan_plan S` may pass
 * `reinterpret_cast<const pars_plan*>(&S)`; include/pars_hip.hpp does exactly
 * that behind a static_assert on the size.
 *
 * Pure C99, no HIP dependency: the CPU restatement under oracle/ includes it too.
 */
#ifndef PARS_PLAN_H
#define PARS_PLAN_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	PARS_MAX_PUNCT		 = 8, /* plan.hpp:19 */
	PARS_MAX_BRACKETS	 = 8, /* plan.hpp:20 */
	PARS_MAX_MODES		 = 8, /* plan.hpp:21 (also the descriptor cap, plan.hpp:75) */
	PARS_MAX_COMMENTS	 = 4, /* plan.hpp:22 */
	PARS_MAX_DFA_STATES	 = 8, /* plan.hpp:52 */
	PARS_MAX_DFA_CLASSES = 8  /* plan.hpp:53 */
};

/* plan.hpp:55 — the factorisation the compiler proved for a descriptor. */
typedef enum {
	PARS_LEX_AFFINE	  = 0,
	PARS_LEX_COMMENT  = 1,
	PARS_LEX_DOUBLED  = 2,
	PARS_LEX_TABLE	  = 3,
	PARS_LEX_FALLBACK = 4
} pars_lex_kind;

/* plan.hpp:24-28 */
typedef struct {
	char open;
	char close;
} pars_bracket_spec;

/* plan.hpp:30-37: escape == 0 none, escape == delim doubled, else prefix escape */
typedef struct {
	char delim;
	char escape;
} pars_mode_spec;

/* plan.hpp:57-73: one lexical masking rule as a DFA over byte classes. */
typedef struct {
	uint8_t kind; /* pars_lex_kind */
	uint8_t n_states;
	uint8_t n_classes;
	uint8_t start;
	uint8_t ok;
	char	delim;
	char	escape;
	uint8_t rep_of[256];								   /* byte -> class */
	uint8_t trans[PARS_MAX_DFA_STATES][PARS_MAX_DFA_CLASSES];  /* state x class -> state */
	uint8_t masked[PARS_MAX_DFA_STATES][PARS_MAX_DFA_CLASSES]; /* state x class -> bit */
} pars_lex_desc;

/* plan.hpp:39-79 */
typedef struct {
	char			  punctuation[PARS_MAX_PUNCT];
	pars_bracket_spec brackets[PARS_MAX_BRACKETS];
	pars_mode_spec	  modes[PARS_MAX_MODES];
	char			  comments[PARS_MAX_COMMENTS];
	uint8_t			  n_punctuation;
	uint8_t			  n_brackets;
	uint8_t			  n_modes;
	uint8_t			  n_comments;
	pars_lex_desc	  descs[PARS_MAX_MODES];
	uint8_t			  n_descs;
} pars_plan;

#define PARS_PLAN_BYTES 3177u

/* Class bits of the R1 classifier (reference: include/pars/model/classify.hpp:13-29). */
#define PARS_CBIT_STRUCTURAL (1u << 0)
#define PARS_CBIT_OPEN (1u << 1)
#define PARS_CBIT_CLOSE (1u << 2)
#define PARS_CBIT_MODE_DELIM(i) (1u << (3 + 2 * (i)))
#define PARS_CBIT_MODE_ESCAPE(i) (1u << (4 + 2 * (i)))
#define PARS_CBIT_COMMENT(i) (1u << (19 + (i)))

#ifdef __cplusplus
}
static_assert(sizeof(pars_plan) == PARS_PLAN_BYTES, "pars_plan must match scan_plan byte for byte");
#endif

#endif /* PARS_PLAN_H */
