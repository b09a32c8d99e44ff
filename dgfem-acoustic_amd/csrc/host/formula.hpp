// Formula sources: evaluation of the expression of a `formula` source at time t, with the
// semantics of the reference's evaluator (src/eqEdit.cpp:107-311, called per step through
// Sources::value, include/configParser.h:32-36):
//  - binary + - (level 2), * / (level 3), ^ (level 4), all left-associative
//    (2^3^2 = 64); unary minus (level 3) only at the start or right after '(' (so 2*-3 is
//    an error); numbers start with a digit (strtod); identifiers are alphanumeric runs:
//    a function when followed by '(', else the constants pi and e, else the variable t;
//  - functions abs int frac rond log(=log10) ln exp sin cos tan sh ch th asin acos atan
//    ash ach ath;
//  - any error evaluates to -1 (the reference prints a message and returns -1).
// Whitespace is removed by the config parsers before expressions get here.
#pragma once
#include <string>

namespace dg {

double evalFormula(const std::string& expr, double t);

}  // namespace dg
