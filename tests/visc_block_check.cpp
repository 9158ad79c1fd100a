// visc_block (the assembly's 3x3 block for corners known at run time) against visc_entry
// (one entry, the element viscosity matrix in the reference's operation order) bit for bit,
// for every corner pair, on random shape gradients and moduli in FP32 and FP64. Both are
// host-and-device functions of kernels.h; tests/test_visc_block.py builds and runs this.
#include "kernels.h"
#include <cstdio>
#include <cstring>
#include <random>
template <typename T> int check(std::mt19937_64 &g) {
    std::uniform_real_distribution<double> u(-2.0, 2.0), w(0.1, 10.0);
    int bad = 0;
    for (int trial = 0; trial < 20000; ++trial) {
        T d[12];
        for (int k = 0; k < 12; ++k) d[k] = (T)u(g);
        d[0] = -(d[1] + d[2] + d[3]); d[4] = -(d[5] + d[6] + d[7]); d[8] = -(d[9] + d[10] + d[11]);
        const T A = (T)w(g), B = (T)w(g), vol = (T)w(g);
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b) {
                T out[9];
                ffk::visc_block<T>(d, A, B, vol, a, b, out);
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) {
                        const T r = ffk::visc_entry<T>(d, A, B, vol, i, a, j, b);
                        bad += std::memcmp(&r, &out[3 * i + j], sizeof(T)) != 0;
                    }
            }
    }
    return bad;
}
int main() {
    std::mt19937_64 g(1);
    const int bf = check<float>(g), bd = check<double>(g);
    std::printf("mismatches: float %d double %d\n", bf, bd);
    return bf != 0 || bd != 0;
}
