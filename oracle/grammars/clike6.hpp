// oracle/grammars/clike6.hpp — the six-mode C-like lexer grammar of the 1 GiB
// C-like workload, written in the reference's grammar DSL (include/pars/dsl.hpp)
// so that the reference compiler (include/pars/compile/plan.hpp:711-748) lowers
// it to a scan_plan.  Compiled only by oracle/ref_driver.cpp, in the container
// that holds the reference; the plan it yields is committed as data
// (pars_amd/plans/clike6.plan).
//
// Modes: code, string literal, char literal, line comment, block comment,
// preprocessor line.  Structural bytes inside `code`: ';' ',' and the bracket
// pairs () [] {} — everything else is a lexical token that never masks.
#pragma once

#include "pars/pars.hpp"

namespace pars_amd::grammars {

using namespace pars;

struct clike6_grammar {
	struct file {};
	struct code {};
	struct ws {};
	struct ident {};
	struct number {};
	struct op {};
	struct str {};
	struct chr {};
	struct line_comment {};
	struct block_comment {};
	struct pp_line {};
	struct paren {};
	struct index {};
	struct block {};

	using rules = type_list<
		file, code, ws, ident, number, op, str, chr, line_comment, block_comment, pp_line, paren,
		index, block>;
	using lexical_rules =
		type_list<ws, ident, number, op, str, chr, line_comment, block_comment, pp_line>;
	using start = file;
	template<class>
	struct def;
};

using G6 = clike6_grammar;

template<>
struct G6::def<G6::ws> {
	using type = star<cls<' ', '\t', '\n', '\r'>>;
};

template<>
struct G6::def<G6::ident> {
	using first = cls<
		'a', 'b', 'c', 'd', 'e', 'f', 'g', 'h', 'i', 'j', 'k', 'l', 'm', 'n', 'o', 'p', 'q', 'r', 's',
		't', 'u', 'v', 'w', 'x', 'y', 'z', 'A', 'B', 'C', 'D', 'E', 'F', 'G', 'H', 'I', 'J', 'K', 'L',
		'M', 'N', 'O', 'P', 'Q', 'R', 'S', 'T', 'U', 'V', 'W', 'X', 'Y', 'Z', '_'>;
	using rest = cls<
		'a', 'b', 'c', 'd', 'e', 'f', 'g', 'h', 'i', 'j', 'k', 'l', 'm', 'n', 'o', 'p', 'q', 'r', 's',
		't', 'u', 'v', 'w', 'x', 'y', 'z', 'A', 'B', 'C', 'D', 'E', 'F', 'G', 'H', 'I', 'J', 'K', 'L',
		'M', 'N', 'O', 'P', 'Q', 'R', 'S', 'T', 'U', 'V', 'W', 'X', 'Y', 'Z', '_', '0', '1', '2', '3',
		'4', '5', '6', '7', '8', '9'>;
	using type = seq<first, star<rest>>;
};

template<>
struct G6::def<G6::number> {
	using digit = cls<'0', '1', '2', '3', '4', '5', '6', '7', '8', '9'>;
	using type	= seq<digit, star<cls<
							   '0', '1', '2', '3', '4', '5', '6', '7', '8', '9', '.', 'x', 'e', 'E',
							   'u', 'U', 'l', 'L', 'f', 'F', 'a', 'b', 'c', 'd'>>>;
};

template<>
struct G6::def<G6::op> {
	using type = plus<cls<'+', '-', '*', '/', '%', '=', '<', '>', '!', '&', '|', '^', '~', '?', ':', '.'>>;
};

template<>
struct G6::def<G6::str> {
	using type =
		seq<lit<'"'>, star<alt<seq<lit<'\\'>, any>, seq<notp<cls<'"', '\\'>>, any>>>, lit<'"'>>;
};

template<>
struct G6::def<G6::chr> {
	using type =
		seq<lit<'\''>, star<alt<seq<lit<'\\'>, any>, seq<notp<cls<'\'', '\\'>>, any>>>, lit<'\''>>;
};

template<>
struct G6::def<G6::line_comment> {
	using type = seq<lit<'/'>, lit<'/'>, star<seq<notp<cls<'\n'>>, any>>>;
};

template<>
struct G6::def<G6::block_comment> {
	using type = seq<
		lit<'/'>, lit<'*'>,
		star<alt<seq<notp<cls<'*'>>, any>, seq<plus<lit<'*'>>, seq<notp<cls<'*', '/'>>, any>>>>,
		plus<lit<'*'>>, lit<'/'>>;
};

template<>
struct G6::def<G6::pp_line> {
	using type = seq<lit<'#'>, star<seq<notp<cls<'\n'>>, any>>>;
};

template<>
struct G6::def<G6::paren> {
	using type = seq<lit<'('>, rule<G6::code>, lit<')'>>;
};

template<>
struct G6::def<G6::index> {
	using type = seq<lit<'['>, rule<G6::code>, lit<']'>>;
};

template<>
struct G6::def<G6::block> {
	using type = seq<lit<'{'>, rule<G6::code>, lit<'}'>>;
};

template<>
struct G6::def<G6::code> {
	using type = star<alt<
		rule<G6::ws>, rule<G6::ident>, rule<G6::number>, rule<G6::op>, rule<G6::str>, rule<G6::chr>,
		rule<G6::line_comment>, rule<G6::block_comment>, rule<G6::pp_line>, lit<';'>, lit<','>,
		rule<G6::paren>, rule<G6::index>, rule<G6::block>>>;
};

template<>
struct G6::def<G6::file> {
	using type = rule<G6::code>;
};

}  // namespace pars_amd::grammars
