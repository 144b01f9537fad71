// The drop-in with OL_PREC_T_FLOAT_COMPAT: PREC_T = float calls use the reference's
// float arithmetic (its tests' default, tests/para.h:8).  Writes <dir>/asm_f32.bin
// (AsmProp of <dir>/field.bin, n x n) and <dir>/cgh_f32.bin (Cgh of <dir>/points.bin,
// m x 3 float, onto n x n), parameters from <dir>/params.txt ("n m").
#define OL_PREC_T_FLOAT_COMPAT
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "openol.h"

typedef float PREC_T;

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    int n = 0, m = 0;
    {
        std::ifstream f(dir + "/params.txt");
        f >> n >> m;
    }
    PREC_T p = 8e-6, lambda = 532e-9, d = 0.05;
    try {
        auto u = std::make_unique<std::complex<PREC_T>[]>((size_t)n * n);
        std::ifstream f(dir + "/field.bin", std::ios::binary);
        f.read(reinterpret_cast<char*>(u.get()), sizeof(std::complex<float>) * (size_t)n * n);
        ol::AsmProp(u, n, n, p, p, lambda, d);
        std::ofstream(dir + "/asm_f32.bin", std::ios::binary)
            .write(reinterpret_cast<const char*>(u.get()), sizeof(std::complex<float>) * (size_t)n * n);
        std::vector<float> xyz((size_t)m * 3);
        std::ifstream g(dir + "/points.bin", std::ios::binary);
        g.read(reinterpret_cast<char*>(xyz.data()), sizeof(float) * xyz.size());
        ol::Object<PREC_T> obj(m);
        for (int j = 0; j < m; j++) obj[j] = ol::Vect3<PREC_T>(xyz[3 * j], xyz[3 * j + 1], xyz[3 * j + 2]);
        auto U = std::make_unique<std::complex<PREC_T>[]>((size_t)n * n);
        ol::Cgh(obj, U, n, n, p, lambda);
        std::ofstream(dir + "/cgh_f32.bin", std::ios::binary)
            .write(reinterpret_cast<const char*>(U.get()), sizeof(std::complex<float>) * (size_t)n * n);
    } catch (const std::exception& e) {
        std::cout << "error: " << e.what() << "\n";
        return 1;
    }
    std::cout << "done\n";
    return 0;
}
