// Driver for the reference formula evaluator (TEST INFRASTRUCTURE ONLY).
// Built by oracle/build_ref.sh together with the reference's src/eqEdit.cpp into
// oracle/_ref/eqedit_ref. Reads lines "<expression>\t<t>" from stdin and prints
// EQ_EDIT::value(true, expression, {{"t", t}}) (the call made by Sources::value,
// include/configParser.h:32-36) with 17 significant digits, one per line.
#include <cstdio>
#include <iostream>
#include <string>
#include <vector>

#include "eqEdit.h"

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        const size_t tab = line.rfind('\t');
        if (tab == std::string::npos) continue;
        const std::string expr = line.substr(0, tab);
        const double t = std::stod(line.substr(tab + 1));
        EQ_EDIT ev;
        std::vector<vartyp> vars = {{"t", t}};
        const double v = ev.value(true, expr, vars);
        std::printf("%.17g\n", v);
        std::fflush(stdout);
    }
    return 0;
}
