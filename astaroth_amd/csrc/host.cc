// host.cc -- host-side part of the C API: parameter derivation, config loading, host meshes,
// host periodic fill and the seeded Fourier-mode initial condition.
//
// ref src/core/update_params/update_params.cc:19-142 + acc-runtime/built-in/variables.h (derived
// grid parameters), src/utils/config_loader.cc:198-315 (.conf parser), src/core/astaroth.cc:295-420
// (host meshes, rand()/RAND_MAX randomisation), src/utils/memory.cc:75-149 (host periodic fill),
// config/astaroth.conf (defaults).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/astaroth.h"

namespace {

struct NamedParam {
    const char* name;
    int kind;  // 0 int, 1 int3, 2 real, 3 real3, 4 bool
    int index;
};

const NamedParam kParams[] = {
    {"AC_max_steps", 0, AC_max_steps},
    {"AC_save_steps", 0, AC_save_steps},
    {"AC_start_step", 0, AC_start_step},
    {"AC_kahan_bits", 0, AC_kahan_bits},
    {"AC_solve_tile_x", 0, AC_solve_tile_x},
    {"AC_solve_tile_y", 0, AC_solve_tile_y},
    {"AC_solve_zchunk", 0, AC_solve_zchunk},
    {"AC_forcing_period_steps", 0, AC_forcing_period_steps},
    {"AC_solve_kernel", 0, AC_solve_kernel},
    {"AC_ngrid", 1, AC_ngrid},
    {"AC_nlocal", 1, AC_nlocal},
    {"AC_mlocal", 1, AC_mlocal},
    {"AC_mgrid", 1, AC_mgrid},
    {"AC_nmin", 1, AC_nmin},
    {"AC_nlocal_max", 1, AC_nlocal_max},
    {"AC_ngrid_max", 1, AC_ngrid_max},
    {"AC_multigpu_offset", 1, AC_multigpu_offset},
    {"AC_cdt", 2, AC_cdt},
    {"AC_cdtv", 2, AC_cdtv},
    {"AC_cdts", 2, AC_cdts},
    {"AC_nu_visc", 2, AC_nu_visc},
    {"AC_cs_sound", 2, AC_cs_sound},
    {"AC_zeta", 2, AC_zeta},
    {"AC_eta", 2, AC_eta},
    {"AC_mu0", 2, AC_mu0},
    {"AC_chi", 2, AC_chi},
    {"AC_cp_sound", 2, AC_cp_sound},
    {"AC_gamma", 2, AC_gamma},
    {"AC_lnT0", 2, AC_lnT0},
    {"AC_lnrho0", 2, AC_lnrho0},
    {"AC_K_heatcond", 2, AC_K_heatcond},
    {"AC_cs2_sound", 2, AC_cs2_sound},
    {"AC_dsmin", 2, AC_dsmin},
    {"AC_current_time", 2, AC_current_time},
    {"AC_dt", 2, AC_dt},
    {"AC_forcing_magnitude", 2, AC_forcing_magnitude},
    {"AC_relhel", 2, AC_relhel},
    {"AC_kmin", 2, AC_kmin},
    {"AC_kmax", 2, AC_kmax},
    {"AC_forcing_phase", 2, AC_forcing_phase},
    {"AC_k_forcex", 2, AC_k_forcex},
    {"AC_k_forcey", 2, AC_k_forcey},
    {"AC_k_forcez", 2, AC_k_forcez},
    {"AC_ff_hel_rex", 2, AC_ff_hel_rex},
    {"AC_ff_hel_rey", 2, AC_ff_hel_rey},
    {"AC_ff_hel_rez", 2, AC_ff_hel_rez},
    {"AC_ff_hel_imx", 2, AC_ff_hel_imx},
    {"AC_ff_hel_imy", 2, AC_ff_hel_imy},
    {"AC_ff_hel_imz", 2, AC_ff_hel_imz},
    {"AC_kaver", 2, AC_kaver},
    {"AC_switch_forcing", 2, AC_switch_forcing},
    {"AC_forcing_period_t", 2, AC_forcing_period_t},
    {"AC_ds", 3, AC_ds},
    {"AC_inv_ds", 3, AC_inv_ds},
    {"AC_inv_ds_2", 3, AC_inv_ds_2},
    {"AC_len", 3, AC_len},
    {"AC_lupwd", 4, AC_lupwd},
    {"AC_lentropy", 4, AC_lentropy},
    {"AC_additive_timestep", 4, AC_additive_timestep},
    {"AC_periodic_grid", 4, AC_periodic_grid_all},
    {"AC_lforcing", 4, AC_lforcing},
};

inline int3 mk3(int x, int y, int z)
{
    int3 v;
    v.x = x;
    v.y = y;
    v.z = z;
    return v;
}

inline size_t mesh_count(const AcMeshInfo& info)
{
    const int3 m = info.int3_params[AC_mlocal];
    return (size_t)m.x * (size_t)m.y * (size_t)m.z;
}

// SplitMix64 (S. Vigna, public domain algorithm).
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Mode {
    int k[3];
    double phase;
    double e[3];
};

std::vector<Mode> draw_modes(uint64_t seed, int nmodes, int kmax)
{
    SplitMix64 rng{seed};
    std::vector<Mode> modes;
    const int span = 2 * kmax + 1;
    for (int m = 0; m < nmodes; ++m) {
        Mode md;
        for (;;) {
            int k2 = 0;
            for (int d = 0; d < 3; ++d) {
                md.k[d] = (int)(rng.uniform() * span) - kmax;
                k2 += md.k[d] * md.k[d];
            }
            if (k2 >= 1 && k2 <= kmax * kmax)
                break;
        }
        md.phase = 2.0 * M_PI * rng.uniform();
        const double kn = std::sqrt((double)(md.k[0] * md.k[0] + md.k[1] * md.k[1] + md.k[2] * md.k[2]));
        for (;;) {
            double r[3], rk = 0.0;
            for (int d = 0; d < 3; ++d) {
                r[d] = 2.0 * rng.uniform() - 1.0;
                rk += r[d] * md.k[d] / kn;
            }
            double nrm = 0.0;
            for (int d = 0; d < 3; ++d) {
                md.e[d] = r[d] - rk * md.k[d] / kn;
                nrm += md.e[d] * md.e[d];
            }
            nrm = std::sqrt(nrm);
            if (nrm > 1e-6) {
                for (int d = 0; d < 3; ++d)
                    md.e[d] /= nrm;
                break;
            }
        }
        modes.push_back(md);
    }
    return modes;
}

template <typename F>
void parallel_for(int n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0)
        nt = 1;
    if (nt > 16)
        nt = 16;
    if ((int)nt > n)
        nt = (unsigned)std::max(n, 1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=]() {
            for (int i = (int)t; i < n; i += (int)nt)
                f(i);
        });
    for (auto& x : th)
        x.join();
}

// Adds sum_m ampl * e_m[c] * cos(k_m . x + phi_m) to three component arrays, evaluated
// separably with complex exponentials per axis.
void add_modes(const AcMeshInfo& info, const std::vector<Mode>& modes, double ampl, double* out[3])
{
    const int3 m = info.int3_params[AC_mlocal];
    const int3 nmin = info.int3_params[AC_nmin];
    const AcReal3 ds = info.real3_params[AC_ds];
    const double dsv[3] = {(double)ds.x, (double)ds.y, (double)ds.z};
    const int mm[3] = {m.x, m.y, m.z};
    const int n0[3] = {nmin.x, nmin.y, nmin.z};
    for (const Mode& md : modes) {
        std::vector<double> cr[3], ci[3];
        for (int d = 0; d < 3; ++d) {
            cr[d].resize(mm[d]);
            ci[d].resize(mm[d]);
            for (int i = 0; i < mm[d]; ++i) {
                const double x = (i - n0[d]) * dsv[d];
                cr[d][i] = std::cos(md.k[d] * x);
                ci[d][i] = std::sin(md.k[d] * x);
            }
        }
        const double pr = std::cos(md.phase), pi = std::sin(md.phase);
        parallel_for(m.z, [&](int k) {
            // e^{i phi} e^{i kz z}
            const double zr = pr * cr[2][k] - pi * ci[2][k];
            const double zi = pr * ci[2][k] + pi * cr[2][k];
            for (int j = 0; j < m.y; ++j) {
                const double yr = zr * cr[1][j] - zi * ci[1][j];
                const double yi = zr * ci[1][j] + zi * cr[1][j];
                const size_t base = (size_t)m.x * (j + (size_t)m.y * k);
                for (int i = 0; i < m.x; ++i) {
                    const double c = yr * cr[0][i] - yi * ci[0][i];
                    for (int q = 0; q < 3; ++q)
                        out[q][base + i] += ampl * md.e[q] * c;
                }
            }
        });
    }
}

}  // namespace

namespace acamd {
// Tables for evaluating the Fourier initial condition on a device exactly as add_modes does on
// the host. Layout (doubles), for group g (0: uu, 1: aa) and mode j of that group:
//   head[(g * nmodes + j) * 5 + {0..4}] = cos(phase), sin(phase), ampl e_x, ampl e_y, ampl e_z
//   axis a in {x, y, z} with extent m_a: cos(k_a x_i), sin(k_a x_i) for i < m_a, stored at
//   tab + ((g * nmodes + j) * (mx + my + mz) + off_a + i) * 2 + {0, 1}, off = (0, mx, mx + my).
void fourier_tables(const AcMeshInfo& info, uint64_t seed, int nmodes, int kmax, double ampl,
                    std::vector<double>& head, std::vector<double>& tab)
{
    const int3 m = info.int3_params[AC_mlocal];
    const int3 nmin = info.int3_params[AC_nmin];
    const AcReal3 ds = info.real3_params[AC_ds];
    const double dsv[3] = {(double)ds.x, (double)ds.y, (double)ds.z};
    const int mm[3] = {m.x, m.y, m.z};
    const int n0[3] = {nmin.x, nmin.y, nmin.z};
    const int off[3] = {0, m.x, m.x + m.y};
    const size_t per = (size_t)(m.x + m.y + m.z);
    head.assign((size_t)2 * nmodes * 5, 0.0);
    tab.assign((size_t)2 * nmodes * per * 2, 0.0);
    for (int g = 0; g < 2; ++g) {
        const std::vector<Mode> modes = draw_modes(seed + (uint64_t)g, nmodes, kmax);
        for (int j = 0; j < nmodes; ++j) {
            const Mode& md = modes[j];
            double* h = &head[((size_t)g * nmodes + j) * 5];
            h[0] = std::cos(md.phase);
            h[1] = std::sin(md.phase);
            for (int q = 0; q < 3; ++q)
                h[2 + q] = ampl * md.e[q];
            double* t = &tab[((size_t)g * nmodes + j) * per * 2];
            for (int d = 0; d < 3; ++d)
                for (int i = 0; i < mm[d]; ++i) {
                    const double x = (i - n0[d]) * dsv[d];
                    t[(off[d] + i) * 2 + 0] = std::cos(md.k[d] * x);
                    t[(off[d] + i) * 2 + 1] = std::sin(md.k[d] * x);
                }
        }
    }
}

uint64_t lnrho_seed(uint64_t seed) { return (seed + 2ULL) * 0xD1B54A32D192ED03ULL; }
}  // namespace acamd

extern "C" {

int acGetRealSize(void) { return (int)sizeof(AcReal); }
size_t acGetMeshInfoSize(void) { return sizeof(AcMeshInfo); }
size_t acGetMeshSize(void) { return sizeof(AcMesh); }

const char* acGetBuildInfo(void)
{
#if AC_DOUBLE_PRECISION
    return "astaroth_amd fp64 gfx950";
#else
    return "astaroth_amd fp32 gfx950";
#endif
}

AcResult acHostUpdateParams(AcMeshInfo* config)
{
    AcMeshInfo& c = *config;
    const int3 ng = c.int3_params[AC_ngrid];
    if (ng.x <= 0 || ng.y <= 0 || ng.z <= 0)
        return AC_FAILURE;
    int3 nl = c.int3_params[AC_nlocal];
    if (nl.x <= 0 || nl.y <= 0 || nl.z <= 0) {
        nl = ng;
        c.int3_params[AC_nlocal] = nl;
    }
    c.int3_params[AC_nmin] = mk3(NGHOST, NGHOST, NGHOST);
    c.int3_params[AC_mlocal] = mk3(nl.x + 2 * NGHOST, nl.y + 2 * NGHOST, nl.z + 2 * NGHOST);
    c.int3_params[AC_mgrid] = mk3(ng.x + 2 * NGHOST, ng.y + 2 * NGHOST, ng.z + 2 * NGHOST);
    c.int3_params[AC_nlocal_max] = mk3(nl.x + NGHOST, nl.y + NGHOST, nl.z + NGHOST);
    c.int3_params[AC_ngrid_max] = mk3(ng.x + NGHOST, ng.y + NGHOST, ng.z + NGHOST);
    // ds from the domain length when it is given (periodic: len / n), ref variables.h:144-152
    const AcReal3 len = c.real3_params[AC_len];
    if (len.x > 0 && len.y > 0 && len.z > 0) {
        AcReal3 ds;
        ds.x = (AcReal)((double)len.x / ng.x);
        ds.y = (AcReal)((double)len.y / ng.y);
        ds.z = (AcReal)((double)len.z / ng.z);
        c.real3_params[AC_ds] = ds;
    }
    const AcReal3 ds = c.real3_params[AC_ds];
    if (!(ds.x > 0 && ds.y > 0 && ds.z > 0))
        return AC_FAILURE;
    AcReal3 inv, inv2;
    inv.x = (AcReal)(1.0 / (double)ds.x);
    inv.y = (AcReal)(1.0 / (double)ds.y);
    inv.z = (AcReal)(1.0 / (double)ds.z);
    inv2.x = inv.x * inv.x;
    inv2.y = inv.y * inv.y;
    inv2.z = inv.z * inv.z;
    c.real3_params[AC_inv_ds] = inv;
    c.real3_params[AC_inv_ds_2] = inv2;
    c.real_params[AC_cs2_sound] = c.real_params[AC_cs_sound] * c.real_params[AC_cs_sound];
    c.real_params[AC_dsmin] = std::min(ds.x, std::min(ds.y, ds.z));
    return AC_SUCCESS;
}

// This is a single-process build: one device holds the whole grid, so the local mesh is the
// grid and each setter sets both extents, as the reference's build without MPI does for
// acSetGridMeshDims (ref src/core/update_params/update_params.cc:173-177, which pushes the grid
// dims into AC_nlocal). A caller that sets only one of them gets a consistent configuration.
AcResult acSetGridMeshDims(const size_t nx, const size_t ny, const size_t nz, AcMeshInfo* info)
{
    info->int3_params[AC_ngrid] = mk3((int)nx, (int)ny, (int)nz);
    info->int3_params[AC_nlocal] = mk3((int)nx, (int)ny, (int)nz);
    return acHostUpdateParams(info);
}

AcResult acSetLocalMeshDims(const size_t nx, const size_t ny, const size_t nz, AcMeshInfo* info)
{
    info->int3_params[AC_nlocal] = mk3((int)nx, (int)ny, (int)nz);
    info->int3_params[AC_ngrid] = mk3((int)nx, (int)ny, (int)nz);
    return acHostUpdateParams(info);
}

AcResult acSetDefaultConfig(AcMeshInfo* config)
{
    AcMeshInfo& c = *config;
    memset(&c, 0, sizeof(c));
    c.int_params[AC_max_steps] = 100;
    c.int_params[AC_save_steps] = 25;
    c.int3_params[AC_ngrid] = mk3(32, 32, 32);
    const AcReal dsd = (AcReal)0.012271846303085129837744700715935558141395192;
    c.real3_params[AC_ds].x = dsd;
    c.real3_params[AC_ds].y = dsd;
    c.real3_params[AC_ds].z = dsd;
    c.real_params[AC_cdt] = (AcReal)0.4;
    c.real_params[AC_cdtv] = (AcReal)0.3;
    c.real_params[AC_cdts] = (AcReal)1.0;
    c.real_params[AC_nu_visc] = (AcReal)2e-5;
    c.real_params[AC_cs_sound] = (AcReal)1.0;
    c.real_params[AC_zeta] = (AcReal)0.01;
    c.real_params[AC_eta] = (AcReal)5e-5;
    c.real_params[AC_mu0] = (AcReal)1.4;
    c.real_params[AC_chi] = (AcReal)0.0001;
    c.real_params[AC_cp_sound] = (AcReal)1.0;
    c.real_params[AC_gamma] = (AcReal)0.5;
    c.real_params[AC_lnT0] = (AcReal)1.2;
    c.real_params[AC_lnrho0] = (AcReal)1.3;
    c.real_params[AC_K_heatcond] = (AcReal)1e-3;
    c.int_params[AC_forcing_period_steps] = 1;
    c.real_params[AC_forcing_magnitude] = (AcReal)1e-5;
    c.real_params[AC_relhel] = (AcReal)0.0;
    c.real_params[AC_kmin] = (AcReal)0.8;
    c.real_params[AC_kmax] = (AcReal)1.2;
    c.bool_params[AC_lupwd] = 1;
    c.bool_params[AC_lentropy] = 1;
    c.bool_params[AC_periodic_grid_all] = 1;
    return acHostUpdateParams(&c);
}

static std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && isspace((unsigned char)s[a]))
        ++a;
    while (b > a && isspace((unsigned char)s[b - 1]))
        --b;
    return s.substr(a, b - a);
}

static bool parse_bool(const std::string& v)
{
    return v == "T" || v == "True" || v == "true" || v == "1";
}

AcResult acLoadConfig(const char* config_path, AcMeshInfo* config)
{
    FILE* fp = fopen(config_path, "r");
    if (!fp) {
        fprintf(stderr, "acLoadConfig: cannot open %s\n", config_path);
        return AC_FAILURE;
    }
    acSetDefaultConfig(config);
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), fp)) > 0)
        text.append(buf, n);
    fclose(fp);
    // strip /* */ and // comments
    std::string clean;
    for (size_t i = 0; i < text.size();) {
        if (text.compare(i, 2, "/*") == 0) {
            const size_t e = text.find("*/", i + 2);
            i = e == std::string::npos ? text.size() : e + 2;
        }
        else if (text.compare(i, 2, "//") == 0) {
            const size_t e = text.find('\n', i);
            i = e == std::string::npos ? text.size() : e;
        }
        else {
            clean.push_back(text[i++]);
        }
    }
    size_t pos = 0;
    AcResult res = AC_SUCCESS;
    bool nlocal_given = false;
    while (pos < clean.size()) {
        size_t e = clean.find('\n', pos);
        if (e == std::string::npos)
            e = clean.size();
        const std::string line = trim(clean.substr(pos, e - pos));
        pos = e + 1;
        const size_t eq = line.find('=');
        if (line.empty() || eq == std::string::npos)
            continue;
        const std::string key = trim(line.substr(0, eq));
        std::string val = trim(line.substr(eq + 1));
        const NamedParam* p = nullptr;
        for (const NamedParam& q : kParams)
            if (key == q.name)
                p = &q;
        if (!p)
            continue;  // parameters of other subsystems are accepted and ignored
        if (p->kind == 1 && p->index == AC_nlocal)
            nlocal_given = true;
        std::vector<std::string> parts;
        if (!val.empty() && (val[0] == '{' || val[0] == '[')) {
            val = val.substr(1, val.find_last_of("}]") - 1);
            size_t s = 0;
            while (s <= val.size()) {
                size_t c = val.find(',', s);
                if (c == std::string::npos)
                    c = val.size();
                parts.push_back(trim(val.substr(s, c - s)));
                s = c + 1;
            }
        }
        else {
            parts.push_back(val);
        }
        switch (p->kind) {
        case 0: config->int_params[p->index] = atoi(parts[0].c_str()); break;
        case 1:
            if (parts.size() < 3) { res = AC_FAILURE; break; }
            config->int3_params[p->index] = mk3(atoi(parts[0].c_str()), atoi(parts[1].c_str()),
                                                atoi(parts[2].c_str()));
            break;
        case 2: config->real_params[p->index] = (AcReal)strtod(parts[0].c_str(), nullptr); break;
        case 3:
            if (parts.size() < 3) { res = AC_FAILURE; break; }
            config->real3_params[p->index].x = (AcReal)strtod(parts[0].c_str(), nullptr);
            config->real3_params[p->index].y = (AcReal)strtod(parts[1].c_str(), nullptr);
            config->real3_params[p->index].z = (AcReal)strtod(parts[2].c_str(), nullptr);
            break;
        case 4: {
            bool all = true;
            for (const std::string& s : parts)
                all = all && parse_bool(s);
            config->bool_params[p->index] = all ? 1 : 0;
            break;
        }
        }
    }
    if (!nlocal_given)  // single device: the local subdomain is the whole grid
        config->int3_params[AC_nlocal] = mk3(0, 0, 0);
    const AcResult r2 = acHostUpdateParams(config);
    return res == AC_SUCCESS ? r2 : res;
}

AcResult acHostMeshCreate(const AcMeshInfo mesh_info, AcMesh* mesh)
{
    mesh->info = mesh_info;
    if (acHostUpdateParams(&mesh->info) != AC_SUCCESS)
        return AC_FAILURE;
    const size_t count = mesh_count(mesh->info);
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        mesh->vertex_buffer[w] = (AcReal*)calloc(count, sizeof(AcReal));
        if (!mesh->vertex_buffer[w])
            return AC_FAILURE;
    }
    return AC_SUCCESS;
}

AcResult acHostMeshDestroy(AcMesh* mesh)
{
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        free(mesh->vertex_buffer[w]);
        mesh->vertex_buffer[w] = nullptr;
    }
    return AC_SUCCESS;
}

AcResult acHostMeshCopyVertexBuffers(const AcMesh src, AcMesh dst)
{
    const size_t count = mesh_count(src.info);
    if (count != mesh_count(dst.info))
        return AC_FAILURE;
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w)
        if (src.vertex_buffer[w] && dst.vertex_buffer[w])
            memcpy(dst.vertex_buffer[w], src.vertex_buffer[w], count * sizeof(AcReal));
    return AC_SUCCESS;
}

AcResult acHostMeshCopy(const AcMesh src, AcMesh* dst)
{
    if (acHostMeshCreate(src.info, dst) != AC_SUCCESS)
        return AC_FAILURE;
    return acHostMeshCopyVertexBuffers(src, *dst);
}

AcResult acHostMeshRandomize(AcMesh* mesh)
{
    const size_t count = mesh_count(mesh->info);
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        if (!mesh->vertex_buffer[w])
            continue;
        for (size_t i = 0; i < count; ++i)
            mesh->vertex_buffer[w][i] = (AcReal)rand() / (AcReal)RAND_MAX;
    }
    return AC_SUCCESS;
}

AcResult acHostMeshApplyPeriodicBounds(AcMesh* mesh)
{
    const AcMeshInfo& info = mesh->info;
    const int3 n = info.int3_params[AC_nlocal];
    const int3 m = info.int3_params[AC_mlocal];
    const int3 n0 = info.int3_params[AC_nmin];
    for (int w = 0; w < NUM_VTXBUF_HANDLES; ++w) {
        AcReal* a = mesh->vertex_buffer[w];
        if (!a)
            continue;
        for (int k = 0; k < m.z; ++k)
            for (int j = 0; j < m.y; ++j)
                for (int i = 0; i < m.x; ++i) {
                    const bool inside = i >= n0.x && i < n0.x + n.x && j >= n0.y &&
                                        j < n0.y + n.y && k >= n0.z && k < n0.z + n.z;
                    if (inside)
                        continue;
                    const int is = ((i - n0.x) % n.x + n.x) % n.x + n0.x;
                    const int js = ((j - n0.y) % n.y + n.y) % n.y + n0.y;
                    const int ks = ((k - n0.z) % n.z + n.z) % n.z + n0.z;
                    a[i + (size_t)m.x * (j + (size_t)m.y * k)] =
                        a[is + (size_t)m.x * (js + (size_t)m.y * ks)];
                }
    }
    return AC_SUCCESS;
}

static const char* const kFieldNames[NUM_VTXBUF_HANDLES] = {
    "VTXBUF_LNRHO", "VTXBUF_UUX", "VTXBUF_UUY", "VTXBUF_UUZ",
    "VTXBUF_AX",    "VTXBUF_AY",  "VTXBUF_AZ",  "VTXBUF_ENTROPY"};

const char* acGetVertexBufferName(const VertexBufferHandle handle)
{
    if ((int)handle < 0 || (int)handle >= NUM_VTXBUF_HANDLES)
        return "";
    return kFieldNames[handle];
}

// Same columns, order and number formatting as the diagnostics file the reference's standalone
// driver appends to (ref samples/standalone_mpi/main.cc:218-391, "timeseries.ts"): header
// "step  t_step  dt  uu_total_*  bb_total_*  vA_total_*  <field>_min  <field>_rms  <field>_max",
// rows "%d %e %e " then "%e %e %e " per group.
AcResult acHostAppendTimeseries(const char* path, const int step, const double t_step,
                                const double dt, const AcDiagnostics* diag, int* found_nan)
{
    if (!path || !diag)
        return AC_FAILURE;
    FILE* fp = std::fopen(path, "a+");
    if (!fp)
        return AC_FAILURE;
    std::fseek(fp, 0, SEEK_END);
    if (std::ftell(fp) == 0) {
        std::fprintf(fp, "step  t_step  dt  ");
        std::fprintf(fp, "uu_total_min  uu_total_rms  uu_total_max  ");
        std::fprintf(fp, "bb_total_min  bb_total_rms  bb_total_max  ");
        std::fprintf(fp, "vA_total_min  vA_total_rms  vA_total_max  ");
        for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i)
            std::fprintf(fp, "%s_min  %s_rms  %s_max  ", kFieldNames[i], kFieldNames[i],
                         kFieldNames[i]);
        std::fprintf(fp, "\n");
    }
    std::fprintf(fp, "%d %e %e ", step, t_step, dt);
    std::fprintf(fp, "%e %e %e ", diag->uu_min, diag->uu_rms, diag->uu_max);
    std::fprintf(fp, "%e %e %e ", diag->bb_min, diag->bb_rms, diag->bb_max);
    std::fprintf(fp, "%e %e %e ", diag->va_min, diag->va_rms, diag->va_max);
    int nan = 0;
    for (int i = 0; i < NUM_VTXBUF_HANDLES; ++i) {
        std::fprintf(fp, "%e %e %e ", diag->min[i], diag->rms[i], diag->max[i]);
        if (std::isnan(diag->min[i]) || std::isnan(diag->rms[i]) || std::isnan(diag->max[i]))
            nan = 1;
    }
    std::fprintf(fp, "\n");
    const bool ok = std::fflush(fp) == 0;
    std::fclose(fp);
    if (found_nan)
        *found_nan = nan;
    return ok ? AC_SUCCESS : AC_FAILURE;
}

// The restart table of the reference's standalone driver (ref
// samples/standalone_mpi/main.cc:186-203): header at step 0 (the file is then truncated), one
// row per snapshot "use_double,mx,my,mz,step_number,modstep,t_step,t_step(exact)" with the
// time printed as %g and as a hexadecimal float (%a), which reads back exactly.
AcResult acHostAppendSnapshotInfo(const char* dir, const int step, const int modstep,
                                  const double t_step, const AcMeshInfo* info)
{
    if (!dir || !info)
        return AC_FAILURE;
    const std::string path = std::string(dir) + "/snapshots_info.csv";
    FILE* fp = std::fopen(path.c_str(), step == 0 ? "w" : "a");
    if (!fp)
        return AC_FAILURE;
    if (step == 0)
        std::fprintf(fp, "use_double,mx,my,mz,step_number,modstep,t_step,t_step(exact)\n");
    const int3 m = info->int3_params[AC_mlocal];
    std::fprintf(fp, "%d,%d,%d,%d,%d,%d,%g,%a\n", (int)(sizeof(AcReal) == 8), m.x, m.y, m.z, step,
                 modstep, t_step, t_step);
    const bool ok = std::fflush(fp) == 0;
    std::fclose(fp);
    return ok ? AC_SUCCESS : AC_FAILURE;
}

// Last row of snapshots_info.csv: the snapshot a restart loads (ref main.cc:579-605).
AcResult acHostReadSnapshotInfo(const char* dir, int* step, int* modstep, double* t_step)
{
    if (!dir)
        return AC_FAILURE;
    const std::string path = std::string(dir) + "/snapshots_info.csv";
    FILE* fp = std::fopen(path.c_str(), "r");
    if (!fp)
        return AC_FAILURE;
    char line[1024];
    bool header = true, found = false;
    int s = 0, ms = 0;
    double t = 0.0;
    while (std::fgets(line, sizeof(line), fp)) {
        if (header) {
            header = false;
            continue;
        }
        int ud, mx, my, mz, s1, ms1;
        double tg;
        char hex[256];
        if (std::sscanf(line, "%d,%d,%d,%d,%d,%d,%lg,%255s", &ud, &mx, &my, &mz, &s1, &ms1, &tg,
                        hex) == 8) {
            s = s1;
            ms = ms1;
            t = std::strtod(hex, nullptr);
            found = true;
        }
    }
    std::fclose(fp);
    if (!found)
        return AC_FAILURE;
    if (step)
        *step = s;
    if (modstep)
        *modstep = ms;
    if (t_step)
        *t_step = t;
    return AC_SUCCESS;
}

AcResult acHostMeshInitFourier(AcMesh* mesh, const unsigned long long seed, const int nmodes,
                               const int kmax, const double ampl, const double lnrho_perturb)
{
    if (nmodes < 0 || kmax < 1)
        return AC_FAILURE;
    const AcMeshInfo& info = mesh->info;
    const size_t count = mesh_count(info);
    std::vector<double> tmp(count * 3);
    for (int group = 0; group < 2; ++group) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        double* out[3] = {tmp.data(), tmp.data() + count, tmp.data() + 2 * count};
        const std::vector<Mode> modes = draw_modes(seed + (uint64_t)group, nmodes, kmax);
        add_modes(info, modes, ampl, out);
        const int base = group == 0 ? VTXBUF_UUX : VTXBUF_AX;
        for (int q = 0; q < 3; ++q)
            for (size_t i = 0; i < count; ++i)
                mesh->vertex_buffer[base + q][i] = (AcReal)out[q][i];
    }
    for (size_t i = 0; i < count; ++i)
        mesh->vertex_buffer[VTXBUF_ENTROPY][i] = (AcReal)0;
    const uint64_t lseed = acamd::lnrho_seed(seed);
    for (size_t i = 0; i < count; ++i) {
        double v = 0.0;
        if (lnrho_perturb != 0.0) {
            const double u = (double)(mix64(lseed + (uint64_t)i) >> 11) * (1.0 / 9007199254740992.0);
            v = lnrho_perturb * (2.0 * u - 1.0);
        }
        mesh->vertex_buffer[VTXBUF_LNRHO][i] = (AcReal)v;
    }
    return acHostMeshApplyPeriodicBounds(mesh);
}

}  // extern "C"
