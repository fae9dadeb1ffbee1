// Seeded synthetic inputs (host side).  The reference hard-codes x* = 1 and b = A*1
// (SparseMatrix.F90:170-172); the seeded workloads below have no reference counterpart and
// are fixed here: SplitMix64 streams (Steele, Lea & Flood, OOPSLA 2014), separable sine
// tables, every operation in a stated order with no fused multiply-add.
#include <cmath>
#include <cstdint>
#include <vector>

#include "common.hpp"

namespace {
constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ULL;

uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
uint64_t next64(uint64_t &s)
{
    s += kGolden;
    return mix64(s);
}
double unit53(uint64_t z) { return (double)(z >> 11) * 0x1.0p-53; }
}  // namespace

extern "C" {

// x*(ix,iy,iz) = sum_{m<8} a_m * S_m0(ix) * S_m1(iy) * S_m2(iz) + U(-0.05, 0.05), with
// S_md(i) = sin(pi*k_md*(i+1)/(n_d+1) + phi_md), a_m ~ U(-1,1), k_md ~ U{1..4},
// phi_md ~ U[0, 2pi).  Random phases keep each term off the matrix's eigenvectors.
int hpccg_xstar_sines(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double *xstar)
{
    HPCCG_CHECK_ARG(nx >= 1 && ny >= 1 && nz >= 1 && xstar, "bad argument");
    const int dims[3] = {nx, ny, nz};
    double amp[8], phi[8][3];
    int kk[8][3];
    uint64_t s = seed;
    for (int m = 0; m < 8; ++m) {
        amp[m] = 2.0 * unit53(next64(s)) - 1.0;
        for (int d = 0; d < 3; ++d) {
            kk[m][d] = 1 + (int)((next64(s) >> 33) % 4u);
            phi[m][d] = 6.283185307179586 * unit53(next64(s));
        }
    }
    std::vector<double> tab[8][3];
    for (int m = 0; m < 8; ++m)
        for (int d = 0; d < 3; ++d) {
            tab[m][d].resize(dims[d]);
            for (int i = 0; i < dims[d]; ++i) {
                double t = 3.141592653589793 * (double)kk[m][d];
                t = t * (double)(i + 1);
                t = t / (double)(dims[d] + 1);
                t = t + phi[m][d];
                tab[m][d][i] = std::sin(t);
            }
        }
    const uint64_t nseed = seed ^ 0xD1B54A32D192ED03ULL;
    for (int iz = 0; iz < nz; ++iz)
        for (int iy = 0; iy < ny; ++iy)
            for (int ix = 0; ix < nx; ++ix) {
                const int64_t v = (int64_t)ix + (int64_t)nx * ((int64_t)iy + (int64_t)ny * iz);
                double acc = 0.0;
                for (int m = 0; m < 8; ++m) {
                    double t = amp[m] * tab[m][0][ix];
                    t = t * tab[m][1][iy];
                    t = t * tab[m][2][iz];
                    acc = acc + t;
                }
                const double noise = 0.1 * unit53(mix64(nseed + (uint64_t)(v + 1) * kGolden)) - 0.05;
                xstar[v] = acc + noise;
            }
    return 0;
}

// x*_i ~ U(-1, 1), counter-based on i.
int hpccg_xstar_uniform(int64_t n, uint64_t seed, double *xstar)
{
    HPCCG_CHECK_ARG(n >= 0 && (n == 0 || xstar), "bad argument");
    for (int64_t i = 0; i < n; ++i) xstar[i] = 2.0 * unit53(mix64(seed + (uint64_t)(i + 1) * kGolden)) - 1.0;
    return 0;
}

// 3*count grid extents drawn from U{lo..hi} (nx, ny, nz of each system in turn).
int hpccg_batch_dims(int32_t count, int32_t lo, int32_t hi, uint64_t seed, int32_t *dims)
{
    HPCCG_CHECK_ARG(count >= 0 && lo >= 1 && hi >= lo && (count == 0 || dims), "bad argument");
    uint64_t s = seed;
    const uint64_t span = (uint64_t)(hi - lo + 1);
    for (int j = 0; j < 3 * count; ++j) dims[j] = lo + (int32_t)((next64(s) >> 33) % span);
    return 0;
}

}  // extern "C"
