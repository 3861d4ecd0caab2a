// forcing.cc -- host side of the helical forcing: the process-wide generator and the random
// forcing parameters, as the reference library exports them (ref include/astaroth_forcing.h,
// include/astaroth_random.h). The device side (the forcing term added to u) is in
// aux_kernels.hip.
//
// The parameters must come out of the same draws, in the same order, from the same
// std::mt19937 and the same distributions as ref src/utils/helical_forcing.cc, so that a run
// seeded like a reference run forces it with the same sequence of modes. That fixes the draw
// order (k, then e by rejection, then the phase), the distributions (integer index into the
// shell population, uniform reals in [-1, 1) and [-pi, pi)), the population order (x fastest)
// and the floating-point expressions whose results are compared (angle tests, f_k).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <map>
#include <random>
#include <utility>
#include <vector>

#include "../../include/astaroth.h"

namespace {

struct D3 {
    double v[3];
};

std::mt19937& generator()
{
    static std::mt19937 g(19937123u);  // ref src/utils/astaroth_random.cc: rng(19937123)
    return g;
}

double draw_real(double lo, double hi)
{
    std::uniform_real_distribution<double> dist(lo, hi);
    return dist(generator());
}

// Uniform integer in [lo, hi): the reference draws from the closed [lo, hi - 1].
int draw_index(int lo, int hi)
{
    std::uniform_int_distribution<int> dist(lo, hi - 1);
    return dist(generator());
}

double dot3(const D3& a, const D3& b) { return a.v[0] * b.v[0] + a.v[1] * b.v[1] + a.v[2] * b.v[2]; }
double len3(const D3& a) { return std::sqrt(dot3(a, a)); }
D3 cross3(const D3& a, const D3& b)
{
    return D3{{a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
               a.v[0] * b.v[1] - a.v[1] * b.v[0]}};
}
double angle3(const D3& a, const D3& b) { return std::acos(dot3(a, b) / (len3(a) * len3(b))); }
bool near(double a, double b) { return std::fabs(a - b) <= 1e-8 + 1e-5 * std::fabs(b); }

// Integer wave vectors with kmin <= |k| <= kmax, z slowest and x fastest (ref
// helical_forcing.cc gen_k), cached per (kmin, kmax). The reference scans the cube
// [-ceil(kmax^3), ceil(kmax^3)]^3; the points it accepts all lie in [-ceil(kmax), ceil(kmax)]^3
// and come in the same order, so the smaller cube gives the same population.
const std::vector<D3>& shell(double kmin, double kmax)
{
    static std::map<std::pair<double, double>, std::vector<D3>> cache;
    std::vector<D3>& pop = cache[std::make_pair(kmin, kmax)];
    if (pop.empty()) {
        const long long c = (long long)std::ceil(kmax);
        for (long long z = -c; z <= c; ++z)
            for (long long y = -c; y <= c; ++y)
                for (long long x = -c; x <= c; ++x) {
                    const D3 k{{(double)x, (double)y, (double)z}};
                    if (len3(k) >= kmin && len3(k) <= kmax)
                        pop.push_back(k);
                }
    }
    return pop;
}

// A unit vector uniform on the sphere: three uniform draws in [-1, 1) until inside the unit
// ball, then normalised.
D3 unit_vector()
{
    D3 e;
    do {
        const double x = draw_real(-1.0, 1.0);
        const double y = draw_real(-1.0, 1.0);
        const double z = draw_real(-1.0, 1.0);
        e = D3{{x, y, z}};
    } while (dot3(e, e) > 1.0);
    const double n = len3(e);
    return D3{{e.v[0] / n, e.v[1] / n, e.v[2] / n}};
}

AcReal3 to_real3(const D3& a) { return AcReal3{(AcReal)a.v[0], (AcReal)a.v[1], (AcReal)a.v[2]}; }

}  // namespace

extern "C" {

void seed_rng(uint32_t seed) { generator().seed(seed); }

AcReal random_uniform_real_01(void)
{
    std::uniform_real_distribution<AcReal> dist(0, 1);
    return dist(generator());
}

ForcingParams generateHelicalForcingParams(const AcReal relhel, const AcReal magnitude,
                                           const AcReal kmin, const AcReal kmax)
{
    ForcingParams p{};
    const std::vector<D3>& pop = shell((double)kmin, (double)kmax);
    if (pop.empty()) {
        fprintf(stderr, "generateHelicalForcingParams: no integer wave vector with %g <= |k| <= %g\n",
                (double)kmin, (double)kmax);
        return p;
    }
    const D3 k = pop[(size_t)draw_index(0, (int)pop.size())];

    // e: a random unit vector, redrawn while (anti)parallel to k
    D3 e = unit_vector();
    while (near(angle3(k, e), 0.0) || near(angle3(k, e), M_PI))
        e = unit_vector();

    // f_k = (|k| (k x e) + i relhel k x (k x e)) / D
    const double h = (double)relhel;
    const double kk = dot3(k, k), ke = dot3(k, e);
    const double den = std::sqrt(1 + h * h) * kk * std::sqrt(1 - (ke * ke) / kk);
    const D3 kxe = cross3(k, e);
    const D3 kxkxe = cross3(k, kxe);
    const double kl = len3(k);
    D3 fre, fim;
    for (int i = 0; i < 3; ++i) {
        fre.v[i] = (kl * kxe.v[i]) / den;
        fim.v[i] = (h * kxkxe.v[i]) / den;
    }
    const double phase = draw_real(-M_PI, M_PI);

    p.magnitude = magnitude;
    p.k_force = to_real3(k);
    p.ff_hel_re = to_real3(fre);
    p.ff_hel_im = to_real3(fim);
    p.phase = (AcReal)phase;
    p.kaver = (kmax - kmin) / 2 + kmin;
    return p;
}

int loadForcingParamsToMeshInfo(const ForcingParams fp, AcMeshInfo* info)
{
    AcReal* r = info->real_params;
    r[AC_forcing_magnitude] = fp.magnitude;
    r[AC_forcing_phase] = fp.phase;
    r[AC_k_forcex] = fp.k_force.x;
    r[AC_k_forcey] = fp.k_force.y;
    r[AC_k_forcez] = fp.k_force.z;
    r[AC_ff_hel_rex] = fp.ff_hel_re.x;
    r[AC_ff_hel_rey] = fp.ff_hel_re.y;
    r[AC_ff_hel_rez] = fp.ff_hel_re.z;
    r[AC_ff_hel_imx] = fp.ff_hel_im.x;
    r[AC_ff_hel_imy] = fp.ff_hel_im.y;
    r[AC_ff_hel_imz] = fp.ff_hel_im.z;
    r[AC_kaver] = fp.kaver;
    return 0;
}

void printForcingParams(const ForcingParams fp)
{
    printf("Forcing parameters:\n magnitude: %lf\n phase: %lf\n k force: %lf\n          %lf\n"
           "          %lf\n ff hel real: %lf\n            : %lf\n            : %lf\n"
           " ff hel imag: %lf\n            : %lf\n            : %lf\n k aver: %lf\n\n",
           (double)fp.magnitude, (double)fp.phase, (double)fp.k_force.x, (double)fp.k_force.y,
           (double)fp.k_force.z, (double)fp.ff_hel_re.x, (double)fp.ff_hel_re.y,
           (double)fp.ff_hel_re.z, (double)fp.ff_hel_im.x, (double)fp.ff_hel_im.y,
           (double)fp.ff_hel_im.z, (double)fp.kaver);
}

}  // extern "C"
