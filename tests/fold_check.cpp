// Host driver of the exact reductions in pgw_fold.h (the code the MFMA
// epilogue, the digit split and the row mixing run on the device), for
// tests/test_fold_extremes.py.  Reads commands from stdin, one per line:
//   F p x0 .. x_{NS-1}   fold_digit_shifts<NL> (NL = 2 for p <= 65536, else 4)
//   R v                  red63_m31(v)
//   D p x                the NL balanced digits of x (digits_biased)
// and prints one result line per command.
#include <iostream>
#include <sstream>
#include <string>

#include "pgw_fold.h"

using namespace pgw;

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string cmd;
        in >> cmd;
        if (cmd == "F") {
            uint64_t p;
            in >> p;
            const int nl = p <= 65536 ? 2 : 4, ns = 2 * nl - 1;
            int32_t x[7] = {0, 0, 0, 0, 0, 0, 0};
            for (int s = 0; s < ns; ++s) {
                long long v;
                in >> v;
                x[s] = (int32_t)v;
            }
            const ModP m = make_modp((uint32_t)p);
            uint32_t w[8];
            uint64_t c = 1;
            for (int s = 0; s < 8; ++s, c = c * 256 % p) w[s] = (uint32_t)c;
            const uint32_t r = nl == 2 ? fold_digit_shifts<2>(x, m, w) : fold_digit_shifts<4>(x, m, w);
            std::cout << r << "\n";
        } else if (cmd == "R") {
            unsigned long long v;
            in >> v;
            std::cout << red63_m31(v) << "\n";
        } else if (cmd == "D") {
            uint64_t p, x;
            in >> p >> x;
            const int nl = p <= 65536 ? 2 : 4;
            const uint64_t u = digits_biased((uint32_t)x, (uint32_t)p, digit_spec(nl));
            for (int j = 0; j < nl; ++j) std::cout << (int)(int8_t)(((u >> (8 * j)) & 0xff) ^ 0x80) << (j + 1 < nl ? " " : "\n");
        } else if (!cmd.empty()) {
            std::cerr << "unknown command " << cmd << "\n";
            return 2;
        }
    }
    return 0;
}
