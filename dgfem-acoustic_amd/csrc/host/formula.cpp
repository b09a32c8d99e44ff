// Operator-precedence evaluation of source formulas. See formula.hpp for the contract.
#include "formula.hpp"

#include <cctype>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace dg {

namespace {

const char* kFunctions[] = {"abs", "int", "frac", "rond", "log", "ln", "exp", "sin", "cos", "tan",
                            "sh",  "ch",  "th",   "asin", "acos", "atan", "ash", "ach", "ath"};
constexpr int kNumFunctions = 19;

double applyFunction(int k, double x) {
    switch (k) {
        case 0: return std::fabs(x);
        case 1: return std::floor(x);
        case 2: return x - std::floor(x);
        case 3: return (x < 0.5 * (std::ceil(x) + std::floor(x))) ? std::floor(x) : std::ceil(x);
        case 4: return std::log10(x);
        case 5: return std::log(x);
        case 6: return std::exp(x);
        case 7: return std::sin(x);
        case 8: return std::cos(x);
        case 9: return std::tan(x);
        case 10: return std::sinh(x);
        case 11: return std::cosh(x);
        case 12: return std::tanh(x);
        case 13: return std::asin(x);
        case 14: return std::acos(x);
        case 15: return std::atan(x);
        case 16: return (x < 0.0) ? -std::log(std::sqrt(x * x + 1.0) - x) : std::log(x + std::sqrt(x * x + 1.0));
        case 17: return std::log(x + std::sqrt(x * x - 1.0));
        case 18: return 0.5 * std::log((1 + x) / (1 - x));
    }
    return 0.0;
}

// Binding level of an operator on the stack ('(' and function markers bind at 1 so that a
// closing parenthesis stops on them; 0 = not an operator).
int level(char op) {
    switch (op) {
        case ')': case '(': case 'f': return 1;
        case '+': case '-': return 2;
        case '*': case '/': case 'm': return 3;
        case '^': return 4;
        default: return 0;
    }
}

struct Machine {
    std::vector<double> vals;
    std::vector<char> ops;  // '(' (incl. the implicit outer one), 'f', 'm', + - * / ^
    std::vector<int> funcs;
    bool ok = true;

    // Pops operators while the incoming operator binds no tighter; stops after removing the
    // '(' or function marker that a ')' closes. Returns true when such a marker was removed.
    bool reduce(int incoming) {
        while (!ops.empty() && incoming <= level(ops.back())) {
            const char op = ops.back();
            ops.pop_back();
            if (op == '(') return true;
            if (op == 'f') {
                if (vals.empty() || funcs.empty()) { ok = false; return true; }
                vals.back() = applyFunction(funcs.back(), vals.back());
                funcs.pop_back();
                return true;
            }
            if (op == 'm') {
                if (vals.empty()) { ok = false; return false; }
                vals.back() = -vals.back();
                continue;
            }
            if (vals.size() < 2) { ok = false; return false; }
            const double b = vals.back();
            vals.pop_back();
            double& a = vals.back();
            switch (op) {
                case '+': a = a + b; break;
                case '-': a = a - b; break;
                case '*': a = a * b; break;
                case '/': a = a / b; break;
                case '^': a = std::pow(a, b); break;
            }
        }
        return false;
    }
};

}  // namespace

double evalFormula(const std::string& expr, double t) {
    if (expr.empty()) return -1.0;
    Machine m;
    m.ops.push_back('(');
    const char* c = expr.c_str();
    bool expectOperand = true, afterOpen = true;
    int endSeen = 0;
    while (endSeen < 2) {
        if (expectOperand) {
            if (*c == '(') {
                m.ops.push_back('(');
                ++c;
                afterOpen = true;
                continue;
            }
            if (*c == '-' && afterOpen) {
                m.ops.push_back('m');
                ++c;
                afterOpen = false;
                continue;
            }
            if (!std::isalnum(static_cast<unsigned char>(*c))) return -1.0;
            if (std::isdigit(static_cast<unsigned char>(*c))) {
                char* end = nullptr;
                m.vals.push_back(std::strtod(c, &end));
                c = end;
                expectOperand = false;
                continue;
            }
            const char* b = c;
            while (std::isalnum(static_cast<unsigned char>(*c))) ++c;
            const std::string name(b, c);
            if (*c == '(') {
                int k = 0;
                while (k < kNumFunctions && name != kFunctions[k]) ++k;
                if (k == kNumFunctions) return -1.0;
                m.funcs.push_back(k);
                m.ops.push_back('f');
                ++c;
                afterOpen = true;
                continue;
            }
            if (name == "pi")
                m.vals.push_back(M_PI);
            else if (name == "e")
                m.vals.push_back(M_E);
            else if (name == "t")
                m.vals.push_back(t);
            else
                return -1.0;
            expectOperand = false;
            continue;
        }
        // operator position; the end of the text acts as a closing parenthesis
        const char op = (*c == 0) ? ')' : *c++;
        const int lv = level(op);
        if (lv == 0 || op == '(') return -1.0;
        const bool closed = m.reduce(lv);
        if (!m.ok) return -1.0;
        if (op != ')') {
            m.ops.push_back(op);
            expectOperand = true;
            afterOpen = false;
        } else {
            // a ')' that closed nothing leaves the machine expecting an operand
            if (!closed) {
                expectOperand = true;
                afterOpen = false;
            }
            if (*c == 0) ++endSeen;
        }
    }
    if (!m.ops.empty() || m.vals.size() != 1 || !m.funcs.empty()) return -1.0;
    return m.vals[0];
}

}  // namespace dg
