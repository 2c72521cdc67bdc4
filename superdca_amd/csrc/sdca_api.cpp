// sdca_api.cpp -- C ABI of libsdca.so: context, reweighting, objective parity entry and
// the batched L-BFGS solve of many loci on one MI355X.
//
// Reference behaviour (paths relative to the SuperDCA repository root):
//   run_plmDCA / plmDCA_solver          SuperDCA/include/plmDCA.hpp:199-437
//   calculate_weights                   SuperDCA/include/plmDCA_optimizer_parameters.hpp:39-98
//   B_eff and lambda                    SuperDCA/include/plmDCA_optimizer_parameters.hpp:257-284
//   objective and gradient              SuperDCA/include/plmDCA_objective.hpp:38-249
//
// Batched L-BFGS.  Every slot of a batch holds one locus.  Its history is m curvature
// pairs s_p = alpha d and y_p = g_new - g_old stored on the device (BF16 under split
// precision, FP32 otherwise), plus two gradient buffers, the search direction and the
// iterate / trial point.  The host keeps the inner products the two-loop recursion
// needs (s.y, y.y, s.g, y.g, g.g; FP64 products computed on the device while the
// vectors stream, or derived from them when a pair is formed) and runs the recursion on
// coefficient vectors; the device then forms d = c_g g + sum_p (c_s s_p + c_y y_p) in
// the same pass that stores the newest pair.  Each round of the batch performs one
// objective evaluation for every active slot: trial point, energies, softmax, gradient,
// dot products, one small device->host copy.  The line search is the More-Thuente
// method, stepped on the host per slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <numeric>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <cstdlib>
#include <deque>
#include <map>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <random>
#include <thread>

#include <sched.h>
#include <string>
#include <vector>

#include "lbfgs_coeffs.hpp"
#include "linesearch.hpp"
#include "sdca_internal.hpp"

namespace sdca {
size_t ref_index(int L, int q, int j, int a, int s);  // sdca_host.cpp
}

namespace {
thread_local std::string g_err;
}

// shared with sdca_warm.cpp
int sdca_fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

namespace {

int fail(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                                      \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess) return fail(SDCA_E_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

template <typename T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SDCA_E_HIP, "hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
    }
    e = hipMemset(*p, 0, count * sizeof(T));
    if (e != hipSuccess) return fail(SDCA_E_HIP, "hipMemset failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

bool valid_history(int m)
{
    return m == 0 || m == 1 || m == 3 || m == 5 || m == 8 || m == 10 || m == 16 || m == 24 || m == 32;
}

// The L-BFGS history of a solve: opts.history, or with 0 the formulation's default (Gram
// 32: its vector passes are cheap against the evaluations a longer history saves; primal
// 10).
int history_for(const sdca_opts &o, bool gram) { return o.history > 0 ? o.history : (gram ? 32 : 10); }

// Loci in flight: opts.batch_loci, or with 0 the formulation's default (Gram slots are
// small -- 6.1 MB at B = 3000, q = 5 -- so 4096 of them; primal slots hold q x K_x vectors).
int slots_for(const sdca_opts &o, bool gram, int K)
{
    const int b = o.batch_loci > 0 ? o.batch_loci : (gram ? 3072 : 128);
    return std::min(b, K);
}

}  // namespace

void sdca_ctx::free_workspace()
{
    dfree(d_gbuf);
    dfree(d_dbuf);
    dfree(d_S);
    dfree(d_Y);
    dfree(d_Xtr);
    d_Xc = nullptr;  // second half of the d_Xtr allocation
    dfree(d_X0);
    dfree(d_Whi);
    dfree(d_Wlo);
    dfree(d_Wl3);
    dfree(d_G32);
    dfree(d_Ghi);
    dfree(d_Glo);
    dfree(d_Gl3);
    dfree(d_hpart);
    dfree(d_Epart);
    dfree(d_upd_part);
    dfree(d_dot_part);
    dfree(d_sm_part);
    dfree(d_res);
    dfree(d_cmds);
    dfree(d_rows);
    dfree(d_src_off);
    dfree(d_score_loci);
    dfree(d_score_ld);
    dfree(d_score_aln);
    dfree(d_aln);
    dfree(d_groups);
    for (void *h : {(void *)h_cmds, (void *)h_res, (void *)h_rows, (void *)h_src_off, (void *)h_score_loci,
                    (void *)h_wl, (void *)h_score_ld, (void *)h_score_aln, (void *)h_groups})
        if (h) (void)hipHostFree(h);
    h_cmds = nullptr;
    h_res = nullptr;
    h_rows = nullptr;
    h_src_off = nullptr;
    h_score_loci = nullptr;
    h_wl = nullptr;
    h_score_ld = nullptr;
    h_score_aln = nullptr;
    h_groups = nullptr;
    max_groups = 0;
    ws_slots = 0;
    ws_m = 0;
    ws_groups = 0;
}

sdca_ctx::~sdca_ctx()
{
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    free_workspace();
    dfree(d_states);
    dfree(d_statesT);
    dfree(d_state_col);
    dfree(d_X);
    dfree(d_Xt);
    dfree(d_w);
    dfree(d_X8);
    dfree(d_cnt);
    dfree(d_SP);
    dfree(d_topk);
    delete mlp;
    delete gram;
    dfree(d_gram);
    dfree(d_w64);
    dfree(d_subQ);
    dfree(d_subQt);
    dfree(d_subKQ);
    dfree(d_subKQt);
    dfree(d_subMinv);
    for (auto &e : ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &g : gev)
        for (auto &e : g)
            if (e) (void)hipEventDestroy(e);
    for (int g = 0; g < sdca::kMaxGroups; ++g)
        if (gstream[g] && gstream[g] != stream) (void)hipStreamDestroy(gstream[g]);
    if (fstream) (void)hipStreamDestroy(fstream);
    if (stream) (void)hipStreamDestroy(stream);
}

sdca::GramWs::~GramWs()
{
    dfree(Vr);
    dfree(aux);
    dfree(Ap);
    dfree(Zr);
    dfree(Wc);
    dfree(Wst);
    dfree(Cst);
    dfree(Kst);
    dfree(d_galn_r);
    dfree(V);
    dfree(Gp);
    dfree(Zp);
    dfree(upd_part);
    dfree(sm_part);
    dfree(dot_part);
    dfree(res);
    dfree(d_cmds);
    dfree(d_galn);
    dfree(d_aln);
    dfree(M32);
    dfree(Mout);
    dfree(Mhi);
    dfree(Mlo);
    dfree(Ml3);
    dfree(mhpart);
    dfree(d_mcmds);
    dfree(d_mgroups);
    dfree(d_moff);
    dfree(d_soff);
    dfree(d_mrow);
    dfree(d_maln);
    dfree(d_sloc);
    dfree(d_sld);
    dfree(d_rows);
    if (h_stage) (void)hipHostFree(h_stage);
    dfree(d_stage);
    if (h_cmds) (void)hipHostFree(h_cmds);
    if (h_res) (void)hipHostFree(h_res);
    if (h_rows) (void)hipHostFree(h_rows);
    dfree(Pend);
    for (int g = 0; g < sdca::kMaxGroups; ++g) {
        if (gexec[g]) (void)hipGraphExecDestroy(gexec[g]);
        if (graph[g]) (void)hipGraphDestroy(graph[g]);
    }
    for (auto &g : ev)
        for (auto &e : g)
            if (e) (void)hipEventDestroy(e);
    for (auto &e : stage_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : pend_free)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : fev)
        if (e) (void)hipEventDestroy(e);
}

namespace {

// ---------------------------------------------------------------- GEMM autotuner
//
// Both contractions walk their output tiles in XCD super-tiles (gm tile rows by 64/gm
// tile columns per group of eight workgroups); the energy contraction also splits K.
// The best shape depends on the batch (M = slots x q), the alignment depth (N = B for
// the energies) and K_x, so the first workspace of a shape times a small set of
// candidates on the real buffers and keeps the fastest.  Choices are cached per
// process and, when SDCA_TUNE_CACHE names a file, across processes (one text line per
// shape: precision Mpad Bp Kxp ksplit kchunk gm_energy gm_grad tn_energy tn_grad).

struct GemmKey {
    int prec, Mpad, Bp;
    int64_t Kxp;
    bool operator<(const GemmKey &o) const
    {
        if (prec != o.prec) return prec < o.prec;
        if (Mpad != o.Mpad) return Mpad < o.Mpad;
        if (Bp != o.Bp) return Bp < o.Bp;
        return Kxp < o.Kxp;
    }
};

struct GemmConfig {
    int ksplit = 1, kchunk = 0, gm_e = 8, gm_g = 8, tn_e = 128, tn_g = 128;
};

std::mutex g_tune_mu;
std::map<GemmKey, GemmConfig> g_tune;
bool g_tune_file_read = false;

void read_tune_file()
{
    g_tune_file_read = true;
    const char *path = std::getenv("SDCA_TUNE_CACHE");
    if (!path || !*path) return;
    FILE *f = std::fopen(path, "r");
    if (!f) return;
    GemmKey k;
    GemmConfig g;
    long long kxp;
    while (std::fscanf(f, "%d %d %d %lld %d %d %d %d %d %d", &k.prec, &k.Mpad, &k.Bp, &kxp, &g.ksplit, &g.kchunk,
                       &g.gm_e, &g.gm_g, &g.tn_e, &g.tn_g) == 10) {
        k.Kxp = kxp;
        const bool tn_ok = (g.tn_e == 128 || g.tn_e == 256) && (g.tn_g == 128 || g.tn_g == 256);
        if (g.ksplit >= 1 && g.kchunk > 0 && (int64_t)g.ksplit * g.kchunk >= k.Kxp && g.gm_e >= 1 && g.gm_e <= 64 &&
            g.gm_g >= 1 && g.gm_g <= 64 && tn_ok)
            g_tune[k] = g;
    }
    std::fclose(f);
}

bool lookup_config(const GemmKey &k, GemmConfig &g)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (!g_tune_file_read) read_tune_file();
    auto it = g_tune.find(k);
    if (it == g_tune.end()) return false;
    g = it->second;
    return true;
}

void store_config(const GemmKey &k, const GemmConfig &g)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune[k] = g;
    const char *path = std::getenv("SDCA_TUNE_CACHE");
    if (!path || !*path) return;
    if (FILE *f = std::fopen(path, "a")) {
        std::fprintf(f, "%d %d %d %lld %d %d %d %d %d %d\n", k.prec, k.Mpad, k.Bp, (long long)k.Kxp, g.ksplit,
                     g.kchunk, g.gm_e, g.gm_g, g.tn_e, g.tn_g);
        std::fclose(f);
    }
}

// Split-K chunk of a reduction of length K for a requested number of slices (whole BK
// steps per slice).
void set_split(const sdca_ctx *c, int64_t K, int64_t ks, GemmConfig &g)
{
    const int BK = c->opts.precision == SDCA_PREC_SPLIT_BF16 ? 64 : 32;
    ks = std::max<int64_t>(1, std::min<int64_t>(ks, K / BK));
    g.kchunk = (int)round_up((K + ks - 1) / ks, BK);
    g.ksplit = (int)((K + g.kchunk - 1) / g.kchunk);
}
void set_split(const sdca_ctx *c, int64_t ks, GemmConfig &g) { set_split(c, c->Kxp, ks, g); }

// Default rule: >= 2048 energy workgroups (128-wide tiles) with the number of
// super-tiles a multiple of the eight XCDs, super-tiles gm tiles high.
GemmConfig default_config(const sdca_ctx *c, int gm, int tn_e = 128)
{
    GemmConfig g;
    g.gm_e = g.gm_g = gm;
    g.tn_e = tn_e;
    const int64_t tiles = (int64_t)(c->Mg / sdca::kTile) * (c->Bp / sdca::kTile);
    int64_t ks = std::max<int64_t>(1, (2048 + tiles - 1) / tiles);
    const sdca::TileMap tm1 = sdca::make_map(c->Mg / sdca::kTile, c->Bp / tn_e, 1, gm);
    const int64_t sup = (int64_t)tm1.MB * tm1.NB;
    while ((ks * sup) % 8 != 0 && ks < 64) ++ks;
    set_split(c, ks, g);
    return g;
}

// Energy candidates: tile widths 128 and 256, super-tile heights 1..8 and half /
// default / double split-K.
std::vector<GemmConfig> tuning_candidates(const sdca_ctx *c)
{
    std::vector<GemmConfig> v;
    const int Mt = c->Mg / sdca::kTile;
    for (int tn : {128, 256})
        for (int gm = 1; gm <= 8; gm *= 2) {
            if (gm > 1 && gm / 2 >= Mt) break;  // taller than the matrix: same as the previous one
            const GemmConfig d = default_config(c, gm, tn);
            for (int64_t ks : {(int64_t)std::max(1, d.ksplit / 2), (int64_t)d.ksplit, (int64_t)d.ksplit * 2}) {
                GemmConfig g = d;
                set_split(c, ks, g);
                bool dup = false;
                for (const auto &o : v) dup |= o.ksplit == g.ksplit && o.gm_e == g.gm_e && o.tn_e == g.tn_e;
                if (!dup) v.push_back(g);
            }
        }
    return v;
}

float elapsed(hipEvent_t a, hipEvent_t b);

// The GEMM group of member a over the commands of parts P (packed rows [P.row[0], rend)):
// the energy contraction against its one-hot X (K = Kxp, split-K ksplit x kchunk) or the
// gradient contraction against X^T (K = Bp).
sdca::GemmGroup make_group(const sdca_ctx *c, int a, const sdca::Parts &P, int rend, bool energy, int ksplit,
                           int kchunk, int gm, int tn)
{
    const sdca_ctx *m = c->members[a];
    sdca::GemmGroup G;
    std::memset(&G, 0, sizeof(G));
    G.aln = a;
    G.q = m->q;
    G.parts = P;
    const int Mt = (rend - P.row[0]) / sdca::kTile;
    if (energy) {
        G.K = (int)m->Kxp;
        G.kchunk = kchunk;
        G.Bbits = m->d_X;
        G.ldbw = m->Kxp / 32;
        G.tm = sdca::make_map(Mt, m->Bp / tn, ksplit, gm);
    } else {
        G.K = m->Bp;
        G.kchunk = m->Bp;
        G.Bbits = m->d_Xt;
        G.ldbw = m->Bp / 32;
        G.tm = sdca::make_map(Mt, (int)(m->Kxp / tn), 1, gm);
    }
    return G;
}

// Lays the groups of a launch out one after the other; returns the launch's workgroups.
int place_groups(sdca::GemmGroup *gs, int ng)
{
    int grid = 0;
    for (int k = 0; k < ng; ++k) {
        gs[k].tile0 = grid;
        grid += sdca::map_grid(gs[k].tm);
    }
    return grid;
}

// Uploads the energy groups (ne) and gradient groups (ngr) of slot group g's round; they
// land at d_groups + g * 2 * max_groups (energy) and + max_groups (gradient).
int upload_groups(sdca_ctx *c, int g, int ne, int ngr, hipStream_t st)
{
    sdca::GemmGroup *h = c->h_groups + (size_t)g * 2 * c->max_groups;
    sdca::GemmGroup *d = c->d_groups + (size_t)g * 2 * c->max_groups;
    HIP_TRY(hipMemcpyAsync(d, h, (size_t)ne * sizeof(sdca::GemmGroup), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(d + c->max_groups, h + c->max_groups, (size_t)ngr * sizeof(sdca::GemmGroup),
                           hipMemcpyHostToDevice, st));
    return SDCA_OK;
}

// The members' descriptors, uploaded before a solve (weights and lambda may have changed).
int upload_alns(sdca_ctx *c, sdca::AlnDev *dst)
{
    std::vector<sdca::AlnDev> v(c->members.size());
    for (size_t a = 0; a < c->members.size(); ++a) {
        const sdca_ctx *m = c->members[a];
        sdca::AlnDev &d = v[a];
        d.X = m->d_X;
        d.Xt = m->d_Xt;
        d.statesT = m->d_statesT;
        d.w = m->d_w;
        d.state_col = m->d_state_col;
        d.B = m->B;
        d.Bp = m->Bp;
        d.L = m->L;
        d.q = m->q;
        d.gap = m->gap;
        d.Kx = (int)m->Kx;
        d.Kxp = (int)m->Kxp;
        d.lamJ2 = (float)(2.0 * m->lambda_J);
        d.lamh2 = 2.0 * m->lambda_h;
    }
    HIP_TRY(hipMemcpy(dst, v.data(), v.size() * sizeof(sdca::AlnDev), hipMemcpyHostToDevice));
    return SDCA_OK;
}
int upload_alns(sdca_ctx *c) { return upload_alns(c, c->d_aln); }

// Times the candidates on the workspace's own buffers (the first slot group, nslots
// slots) with synthetic slot commands (every slot active, writing its first gradient) and
// leaves the buffers zeroed again.  A context solving its own alignment only.
int autotune(sdca_ctx *c, int nslots, const std::vector<GemmConfig> &cand, GemmConfig &best)
{
    const bool split = c->opts.precision == SDCA_PREC_SPLIT_BF16;
    hipStream_t st = c->gstream[0];  // the first group's stream
    std::vector<sdca::SlotCmd> cmds(nslots);
    for (int i = 0; i < nslots; ++i) {
        std::memset(&cmds[i], 0, sizeof(cmds[i]));
        cmds[i].mode = sdca::MODE_INIT;
        cmds[i].locus = i % c->L;
        cmds[i].gt = 0;  // the first gradient buffer
        cmds[i].gk = -1;
        cmds[i].pnew = -1;
        cmds[i].slot = i;
        cmds[i].row = i * c->q;
        cmds[i].level = sdca::LV_SPLIT;
        cmds[i].aln = 0;
        cmds[i].q = c->q;
        cmds[i].Kx = (int)c->Kx;
        cmds[i].Kxp = (int)c->Kxp;
    }
    // every command at the split level: the fine part is empty, the coarse part starts at Mg
    const sdca::Parts parts{{0, 0, c->Mg}, {0, 0, nslots, nslots}};
    const size_t nw = (size_t)c->Mpad * c->Kxp;
    HIP_TRY(hipMemsetD32Async((hipDeviceptr_t)c->d_Xtr, 0x3c23d70au, nw, st));  // 0.01f
    if (split) {
        HIP_TRY(hipMemsetD16Async((hipDeviceptr_t)c->d_Whi, 0x3c24u, nw, st));  // bf16 0.01
        HIP_TRY(hipMemsetD16Async((hipDeviceptr_t)c->d_Wlo, 0x3700u, nw, st));
    }
    const int reps = 3;
    sdca::GemmGroup *hg = c->h_groups, *dg = c->d_groups;
    int grid_e = 0, grid_g = 0;
    auto set_energy = [&](const GemmConfig &g) -> int {
        for (int i = 0; i < nslots; ++i) cmds[i].ksplit = g.ksplit;
        HIP_TRY(hipStreamSynchronize(st));  // the pinned staging is rewritten
        HIP_TRY(hipMemcpyAsync(c->d_cmds, cmds.data(), nslots * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st));
        hg[0] = make_group(c, 0, parts, c->Mg, true, g.ksplit, g.kchunk, g.gm_e, g.tn_e);
        grid_e = place_groups(hg, 1);
        return upload_groups(c, 0, 1, 0, st);
    };
    auto set_grad = [&](int gm, int tn) -> int {
        HIP_TRY(hipStreamSynchronize(st));
        hg[c->max_groups] = make_group(c, 0, parts, c->Mg, false, 1, 0, gm, tn);
        grid_g = place_groups(hg + c->max_groups, 1);
        return upload_groups(c, 0, 0, 1, st);
    };
    const int64_t ldx = c->ws_ldx, ldb = c->ws_ldb;
    auto energy = [&](const GemmConfig &g) {
        sdca::kern::gemm_energy(st, c->opts.precision, dg, 1, grid_e, g.tn_e, c->d_Xtr, c->d_Whi, c->d_Wlo, c->d_Wl3,
                                ldx, ldb, (int64_t)c->Mg * ldb, c->d_Epart);
        sdca::kern::softmax(st, c->opts.precision, c->d_Epart, (int64_t)c->Mg * ldb, ldb, nslots, c->nbchunk,
                            c->d_cmds, c->d_aln, c->d_G32, c->d_Ghi, c->d_Glo, c->d_Gl3, c->d_sm_part, c->d_hpart);
    };
    auto grad = [&](int tn) {
        sdca::kern::gemm_grad(st, c->opts.precision, dg + c->max_groups, 1, grid_g, tn, c->d_G32, c->d_Ghi,
                              c->d_Glo, c->d_Gl3, ldb, c->d_aln, c->d_cmds, c->d_Xtr, ldx, c->d_gbuf, c->n_slot,
                              c->d_hpart, c->nbchunk);
    };
    auto time = [&](auto &&launch) -> float {
        launch();
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(c->ev[6], st));
        for (int r = 0; r < reps; ++r) launch();
        HIP_TRY(hipEventRecord(c->ev[7], st));
        HIP_TRY(hipEventSynchronize(c->ev[7]));
        return elapsed(c->ev[6], c->ev[7]) / reps;
    };
    int rc;
    float bt = 1e30f;
    for (const auto &g : cand) {
        if ((rc = set_energy(g))) return rc;
        const float t = time([&] { energy(g); });
        if (t < 0) return SDCA_E_HIP;
        if (c->opts.verbose)
            std::fprintf(stderr, "sdca tune: energy+softmax tn %d ksplit %d gm %d: %.3f ms\n", g.tn_e, g.ksplit,
                         g.gm_e, t);
        if (t < bt) {
            bt = t;
            best.ksplit = g.ksplit;
            best.kchunk = g.kchunk;
            best.gm_e = g.gm_e;
            best.tn_e = g.tn_e;
        }
    }
    if ((rc = set_energy(best))) return rc;
    energy(best);  // node beliefs G for the gradient candidates
    bt = 1e30f;
    const int Mt = c->Mg / sdca::kTile;
    for (int tn : {128, 256})
        for (int gm = 1; gm <= 8; gm *= 2) {
            if (gm > 1 && gm / 2 >= Mt) break;
            if ((rc = set_grad(gm, tn))) return rc;
            const float t = time([&] { grad(tn); });
            if (c->opts.verbose) std::fprintf(stderr, "sdca tune: grad tn %d gm %d: %.3f ms\n", tn, gm, t);
            if (t < bt) {
                bt = t;
                best.gm_g = gm;
                best.tn_g = tn;
            }
        }
    // the solve starts from zeroed buffers
    const size_t all = (size_t)c->ngroups * nslots;
    HIP_TRY(hipMemsetAsync(c->d_gbuf, 0, all * 2 * c->n_slot * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(c->d_Xtr, 0, nw * sizeof(float), st));
    if (split) {
        HIP_TRY(hipMemsetAsync(c->d_Whi, 0, nw * sizeof(uint16_t), st));
        HIP_TRY(hipMemsetAsync(c->d_Wlo, 0, nw * sizeof(uint16_t), st));
    }
    HIP_TRY(hipMemsetAsync(c->d_cmds, 0, all * sizeof(sdca::SlotCmd), st));
    HIP_TRY(hipStreamSynchronize(st));
    return SDCA_OK;
}

// Bytes per element of a stored curvature pair: BF16 under split precision (s and y
// are differences, accurate to their own magnitude after rounding; every product uses
// the rounded stored vector), FP32 otherwise.
inline size_t pair_bytes(const sdca_ctx *c) { return c->opts.precision == SDCA_PREC_SPLIT_BF16 ? 2 : 4; }

// Slot groups: with pipeline_groups = 2 the slots split into two halves of whole
// 128-row GEMM tiles whose rounds run on separate streams (the host steps one group while
// the other computes); otherwise one group.  One group is the default: the two groups'
// kernels compete for the same CUs, and at the C4 shape one group is faster.
int choose_groups(const sdca_ctx *c, int nslots)
{
    if (c->opts.pipeline_groups < 2 || c->members.size() != 1) return 1;  // (4 means 2 here)
    const bool even = nslots % 2 == 0 && ((int64_t)(nslots / 2) * c->q) % sdca::kTile == 0;
    return even ? 2 : 1;
}

// A context solving its own alignment only (not the internal context of a multi-alignment
// solve): its workspace dimensions are its own, its GEMM shapes can be tuned.
bool solves_itself(const sdca_ctx *c) { return c->members.size() == 1 && c->members[0] == c; }

// Allocates the batch workspace for nslots concurrent loci and history m, sized for the
// largest member alignment.
int ensure_workspace(sdca_ctx *c, int nslots, int m)
{
    const int ng = choose_groups(c, nslots);
    int wq = 0, wldb = 0, wL = 0;
    int64_t wldx = 0, nsl = 0;
    for (const sdca_ctx *a : c->members) {
        wq = std::max(wq, a->q);
        wldb = std::max(wldb, a->Bp);
        wL = std::max(wL, a->L);
        wldx = std::max(wldx, a->Kxp);
        nsl = std::max(nsl, (int64_t)a->q * a->Kxp);
    }
    const int nmem = (int)c->members.size();
    if (c->ws_slots == nslots && c->ws_m == m && c->ws_groups == ng && c->n_slot == nsl && c->ws_q == wq &&
        c->ws_ldx == wldx && c->ws_ldb == wldb && c->ws_L == wL && c->max_groups == nmem)
        return SDCA_OK;
    c->free_workspace();
    const bool own = solves_itself(c);
    c->n_slot = nsl;
    c->ws_q = wq;
    c->ws_ldx = wldx;
    c->ws_ldb = wldb;
    c->ws_L = wL;
    c->max_groups = nmem;
    c->ngroups = ng;
    c->gslots = nslots / ng;
    c->Mg = (int)round_up((int64_t)c->gslots * wq, sdca::kTile);
    c->Mpad = ng * c->Mg;
    // each (alignment, level) part of a round's packed rows starts on a tile boundary
    c->Mgp = c->Mg + (own ? 2 : 3 * std::min(nmem, c->gslots)) * sdca::kTile;
    const size_t prow = (size_t)ng * c->Mgp;  // packed rows of all groups
    const bool split = c->opts.precision == SDCA_PREC_SPLIT_BF16;
    int rc;
    c->ws_m = m;  // ring shapes below and the autotuner use it
    if ((rc = dalloc(&c->d_gbuf, (size_t)nslots * 2 * c->n_slot))) return rc;
    if ((rc = dalloc(&c->d_dbuf, (size_t)nslots * c->n_slot))) return rc;
    if ((rc = dalloc((uint8_t **)&c->d_S, (size_t)nslots * m * c->n_slot * pair_bytes(c)))) return rc;
    if ((rc = dalloc((uint8_t **)&c->d_Y, (size_t)nslots * m * c->n_slot * pair_bytes(c)))) return rc;
    c->xc_off = (int64_t)prow * wldx;
    if ((rc = dalloc(&c->d_Xtr, prow * wldx + (size_t)nslots * c->n_slot))) return rc;
    c->d_Xc = c->d_Xtr + c->xc_off;
    if (split) {
        if ((rc = dalloc(&c->d_Whi, prow * wldx))) return rc;
        if ((rc = dalloc(&c->d_Wlo, prow * wldx))) return rc;
        if ((rc = dalloc(&c->d_Ghi, prow * wldb))) return rc;
        if ((rc = dalloc(&c->d_Glo, prow * wldb))) return rc;
        if ((rc = dalloc(&c->d_Gl3, prow * wldb))) return rc;
        // the third parts are read only for fine commands: allocated when a coupling
        // threshold can bring a slot to the fine level
        if (c->opts.j_threshold > 0.0 && c->opts.fine_factor > 0 && (rc = dalloc(&c->d_Wl3, prow * wldx)))
            return rc;
    } else {
        if ((rc = dalloc(&c->d_G32, prow * wldb))) return rc;
    }
    // GEMM configuration: the cached/tuned one for this shape, else the default rule; the
    // groups of a multi-alignment solve take 256-wide tiles and a split-K chosen per round
    const GemmKey key{c->opts.precision, c->Mg, c->Bp, c->Kxp};
    const sdca_opts &o = c->opts;
    const bool forced = o.gemm_ksplit > 0 || o.gemm_gm_energy > 0 || o.gemm_gm_grad > 0 || o.gemm_tn_energy > 0 ||
                        o.gemm_tn_grad > 0;
    GemmConfig cfg;
    bool have = own && !forced && lookup_config(key, cfg);
    const bool tune = own && !forced && !have && o.autotune;
    if (!own) {
        cfg.ksplit = 16;
        cfg.kchunk = 0;
        cfg.gm_e = cfg.gm_g = 8;
        cfg.tn_e = cfg.tn_g = 256;
    } else {
        if (!have)
            cfg = default_config(c, o.gemm_gm_energy > 0 ? o.gemm_gm_energy : 8,
                                 o.gemm_tn_energy > 0 ? o.gemm_tn_energy : 128);
        if (forced) {
            if (o.gemm_ksplit > 0) set_split(c, o.gemm_ksplit, cfg);
            cfg.gm_g = o.gemm_gm_grad > 0 ? o.gemm_gm_grad : 8;
            cfg.tn_g = o.gemm_tn_grad > 0 ? o.gemm_tn_grad : 128;
        }
    }
    int max_ksplit = cfg.ksplit;
    std::vector<GemmConfig> cand;
    if (tune) {
        cand = tuning_candidates(c);
        for (const auto &k : cand) max_ksplit = std::max(max_ksplit, k.ksplit);
    }
    if ((rc = dalloc(&c->d_Epart, (size_t)max_ksplit * prow * wldb))) return rc;  // per group: ksplit x Mgp x ldb
    c->ksplit_alloc = max_ksplit;
    // blocks per slot of the vector passes: up to 128 (at C4, 128 took k_update from
    // 3.15 to 3.03 ms per round against 64; 32 was slower)
    c->nchunk = (int)std::min<int64_t>(128, std::max<int64_t>(1, (c->n_slot / 4 + 255) / 256));
    c->nbchunk = (wldb + 255) / 256;
    if ((rc = dalloc(&c->d_hpart, (size_t)nslots * c->nbchunk * 8))) return rc;
    if ((rc = dalloc(&c->d_upd_part, (size_t)nslots * c->nchunk * sdca::ResLayout{m}.nupd()))) return rc;
    if ((rc = dalloc(&c->d_dot_part, (size_t)nslots * c->nchunk * sdca::ResLayout{m}.ndots()))) return rc;
    if ((rc = dalloc(&c->d_sm_part, (size_t)nslots * c->nbchunk))) return rc;
    if ((rc = dalloc(&c->d_res, (size_t)nslots * sdca::ResLayout{m}.width()))) return rc;
    if ((rc = dalloc(&c->d_cmds, (size_t)nslots))) return rc;
    if ((rc = dalloc(&c->d_rows, (size_t)nslots * wL))) return rc;
    if ((rc = dalloc(&c->d_src_off, (size_t)nslots))) return rc;
    if ((rc = dalloc(&c->d_score_loci, (size_t)nslots))) return rc;
    if ((rc = dalloc(&c->d_score_ld, (size_t)nslots))) return rc;
    if ((rc = dalloc(&c->d_score_aln, (size_t)nslots))) return rc;
    if ((rc = dalloc(&c->d_aln, (size_t)nmem))) return rc;
    if ((rc = dalloc(&c->d_groups, (size_t)ng * 2 * nmem))) return rc;
    HIP_TRY(hipHostMalloc((void **)&c->h_cmds, (size_t)nslots * sizeof(sdca::SlotCmd), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_res, (size_t)nslots * sdca::ResLayout{m}.width() * sizeof(double),
                          hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_rows, (size_t)nslots * wL * sizeof(float), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_src_off, (size_t)nslots * sizeof(int64_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_score_loci, (size_t)nslots * sizeof(int32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_score_ld, (size_t)nslots * sizeof(int32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_score_aln, (size_t)nslots * sizeof(int32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_wl, (size_t)3 * nslots * sizeof(int32_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&c->h_groups, (size_t)ng * 2 * nmem * sizeof(sdca::GemmGroup),
                          hipHostMallocDefault));
    c->h_xt_off.assign(nslots, -1);
    // the zero fills above must be complete before any kernel of the solve runs
    HIP_TRY(hipDeviceSynchronize());
    if ((rc = upload_alns(c))) return rc;
    c->tuned = 0;
    if (tune && (rc = autotune(c, c->gslots, cand, cfg))) return rc;
    c->ksplit = cfg.ksplit;
    c->kchunk = cfg.kchunk;
    c->gm_energy = cfg.gm_e;
    c->gm_grad = cfg.gm_g;
    c->tn_energy = cfg.tn_e;
    c->tn_grad = cfg.tn_g;
    c->tuned = tune || have;
    if (tune) store_config(key, cfg);
    c->ws_slots = nslots;
    c->ws_m = m;
    c->ws_groups = ng;
    return SDCA_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

// One batched evaluation round of slot group g (slots [g*gslots, (g+1)*gslots)) on the
// group's stream: trial points, energies, softmax, gradient, dots, and the copy of the
// result rows (ResLayout) to c->h_res.  cmds points at the group's commands.
int launch_round(sdca_ctx *c, int g, const sdca::SlotCmd *cmds)
{
    const int m = c->ws_m, W = sdca::ResLayout{m}.width();
    const bool split = c->opts.precision == SDCA_PREC_SPLIT_BF16;
    const int ns = c->gslots, s0 = g * ns;
    const int nmem = (int)c->members.size();
    const bool own = solves_itself(c);
    // command list: the active slots by alignment (member order), within an alignment by
    // level, finest first; each (alignment, level) part in slot order and starting on a tile
    // boundary of the packed rows.  One GEMM group per alignment with commands.
    std::vector<int> &lst = c->gcmd_slot[g];
    lst.clear();
    sdca::SlotCmd *hc = c->h_cmds + s0;
    std::vector<sdca::Parts> gparts;
    std::vector<int> galn, gend;
    int row = 0;
    for (int a = 0; a < nmem; ++a) {
        const sdca_ctx *A = c->members[a];
        bool any = false;
        for (int i = 0; i < ns && !any; ++i) any = cmds[i].mode != sdca::MODE_IDLE && cmds[i].aln == a;
        if (!any) continue;
        sdca::Parts parts;
        for (int part = 0; part < 3; ++part) {
            const int level = sdca::LV_FINE - part;
            row = (int)round_up(row, sdca::kTile);
            parts.row[part] = row;
            parts.cmd[part] = (int)lst.size();
            for (int i = 0; i < ns; ++i) {
                if (cmds[i].mode == sdca::MODE_IDLE || cmds[i].aln != a || cmds[i].level != level) continue;
                sdca::SlotCmd &k = hc[lst.size()];
                k = cmds[i];
                k.slot = i;
                k.row = row;
                k.q = A->q;
                k.Kx = (int)A->Kx;
                k.Kxp = (int)A->Kxp;
                row += A->q;
                lst.push_back(i);
            }
        }
        parts.cmd[3] = (int)lst.size();
        row = (int)round_up(row, sdca::kTile);
        gparts.push_back(parts);
        galn.push_back(a);
        gend.push_back(row);
    }
    const int nc = (int)lst.size();
    if (nc == 0) return SDCA_OK;
    const int Mrows = row;  // packed rows of all parts
    const int ngr = (int)gparts.size();
    // split-K of the energy contractions.  One alignment: the tuned slices, and a round with
    // few active rows (the end of a solve) takes proportionally more, packed Mrows apart in
    // the group's partial-energy buffer, so that its energy launch still fills the GPU.
    // Several alignments: the same count of slices for every group, about 2048 workgroups.
    int ksplit = c->ksplit, kchunk = c->kchunk, Mz = c->Mgp;
    const int64_t cap = std::min<int64_t>(128, (int64_t)c->ksplit_alloc * c->Mgp / Mrows);
    if (own) {
        if (2 * Mrows <= c->Mg) {
            GemmConfig gs;
            set_split(c, std::min<int64_t>(cap, (int64_t)c->ksplit * c->Mg / Mrows), gs);
            if (gs.ksplit <= cap && gs.ksplit > ksplit) {
                ksplit = gs.ksplit;
                kchunk = gs.kchunk;
                Mz = Mrows;
            }
        }
    } else {
        int64_t tiles = 0;
        for (int k = 0; k < ngr; ++k)
            tiles += (int64_t)(gend[k] - gparts[k].row[0]) / sdca::kTile * (c->members[galn[k]]->Bp / c->tn_energy);
        ksplit = (int)std::max<int64_t>(1, std::min<int64_t>(cap, (2048 + tiles - 1) / tiles));
        Mz = Mrows;
    }
    sdca::GemmGroup *hge = c->h_groups + (size_t)g * 2 * c->max_groups, *hgg = hge + c->max_groups;
    for (int k = 0; k < ngr; ++k) {
        int ks = ksplit, kc = kchunk;
        if (!own) {
            GemmConfig gs;
            set_split(c, c->members[galn[k]]->Kxp, ksplit, gs);
            ks = gs.ksplit;
            kc = gs.kchunk;
        }
        for (int i = gparts[k].cmd[0]; i < gparts[k].cmd[3]; ++i) hc[i].ksplit = ks;
        hge[k] = make_group(c, galn[k], gparts[k], gend[k], true, ks, kc, c->gm_energy, c->tn_energy);
        hgg[k] = make_group(c, galn[k], gparts[k], gend[k], false, 1, 0, c->gm_grad, c->tn_grad);
    }
    const int grid_e = place_groups(hge, ngr), grid_g = place_groups(hgg, ngr);
    const size_t vo = (size_t)s0 * c->n_slot;              // slot-vector offset of the group
    const size_t ro = (size_t)g * c->Mgp * c->ws_ldx;      // its packed trial-point rows
    const size_t go = (size_t)g * c->Mgp * c->ws_ldb;      // its packed node-belief rows
    float *gbuf = c->d_gbuf + (size_t)s0 * 2 * c->n_slot;
    void *dbuf = split ? (void *)(reinterpret_cast<uint16_t *>(c->d_dbuf) + vo) : (void *)(c->d_dbuf + vo);
    const size_t po = (size_t)s0 * m * c->n_slot * pair_bytes(c);
    void *Sp = (uint8_t *)c->d_S + po, *Yp = (uint8_t *)c->d_Y + po;
    sdca::SlotCmd *dc = c->d_cmds + s0;
    float *Epart = c->d_Epart + (size_t)g * c->ksplit_alloc * c->Mgp * c->ws_ldb;
    const sdca::GemmGroup *dge = c->d_groups + (size_t)g * 2 * c->max_groups, *dgg = dge + c->max_groups;
    hipStream_t st = c->gstream[g];
    hipEvent_t *ev = c->gev[g];
    HIP_TRY(hipMemcpyAsync(dc, hc, nc * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st));
    int rc;
    if ((rc = upload_groups(c, g, ngr, ngr, st))) return rc;
    HIP_TRY(hipEventRecord(ev[0], st));
    double *upd_part = c->d_upd_part + (size_t)s0 * c->nchunk * sdca::ResLayout{m}.nupd();
    sdca::kern::update(st, m, split, dc, nc, gbuf, dbuf, Sp, Yp, c->d_Xc + vo, c->d_Xtr + ro,
                       split ? c->d_Whi + ro : nullptr, split ? c->d_Wlo + ro : nullptr,
                       split ? c->d_Wl3 + ro : nullptr, c->d_X0 ? c->d_X0 + vo : nullptr, c->n_slot, c->ws_ldx,
                       upd_part, c->nchunk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[1], st));
    // fine rows: hi + mid + lo, split rows: hi + lo, coarse rows: the high parts (or FP32 throughout)
    const int64_t zstride = (int64_t)Mz * c->ws_ldb;
    sdca::kern::gemm_energy(st, c->opts.precision, dge, ngr, grid_e, c->tn_energy, c->d_Xtr + ro,
                            split ? c->d_Whi + ro : nullptr, split ? c->d_Wlo + ro : nullptr,
                            split ? c->d_Wl3 + ro : nullptr, c->ws_ldx, c->ws_ldb, zstride, Epart);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[2], st));
    double *hpart = c->d_hpart + (size_t)s0 * c->nbchunk * 8;
    sdca::kern::softmax(st, c->opts.precision, Epart, zstride, c->ws_ldb, nc, c->nbchunk, dc, c->d_aln,
                        split ? nullptr : c->d_G32 + go, split ? c->d_Ghi + go : nullptr,
                        split ? c->d_Glo + go : nullptr, split ? c->d_Gl3 + go : nullptr,
                        c->d_sm_part + (size_t)s0 * c->nbchunk, hpart);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[3], st));
    sdca::kern::gemm_grad(st, c->opts.precision, dgg, ngr, grid_g, c->tn_grad, split ? nullptr : c->d_G32 + go,
                          split ? c->d_Ghi + go : nullptr, split ? c->d_Glo + go : nullptr,
                          split ? c->d_Gl3 + go : nullptr, c->ws_ldb, c->d_aln, dc, c->d_Xtr + ro, c->ws_ldx, gbuf,
                          c->n_slot, hpart, c->nbchunk);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[4], st));
    double *dot_part = c->d_dot_part + (size_t)s0 * c->nchunk * sdca::ResLayout{m}.ndots();
    sdca::kern::dots(st, m, split, dc, nc, gbuf, dbuf, Sp, Yp, c->n_slot, dot_part, c->nchunk);
    HIP_TRY(hipGetLastError());
    sdca::kern::reduce_results(st, m, nc, upd_part, dot_part,
                               c->d_sm_part + (size_t)s0 * c->nbchunk, c->nchunk, c->nbchunk,
                               c->d_res + (size_t)s0 * W);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[5], st));
    HIP_TRY(hipMemcpyAsync(c->h_res + (size_t)s0 * W, c->d_res + (size_t)s0 * W, (size_t)nc * W * sizeof(double),
                           hipMemcpyDeviceToHost, st));
    return SDCA_OK;
}

// Waits for group g's round and accounts its kernel times and algorithmic traffic.  The
// result rows arrive in command order (full-precision commands first) and are moved to their
// slots; each listed slot's trial point is recorded for scoring.
int finish_round(sdca_ctx *c, int g, const sdca::SlotCmd *cmds)
{
    const int m = c->ws_m;
    const bool split = c->opts.precision == SDCA_PREC_SPLIT_BF16;
    const std::vector<int> &lst = c->gcmd_slot[g];
    if (lst.empty()) return SDCA_OK;
    HIP_TRY(hipStreamSynchronize(c->gstream[g]));
    {
        const int W = sdca::ResLayout{m}.width(), s0 = g * c->gslots;
        double *res = c->h_res + (size_t)s0 * W;
        std::vector<double> &tmp = c->h_res_cmd;
        tmp.assign(res, res + lst.size() * W);
        for (size_t k = 0; k < lst.size(); ++k) {
            std::memcpy(res + (size_t)lst[k] * W, tmp.data() + k * W, W * sizeof(double));
            c->h_xt_off[s0 + lst[k]] = (int64_t)((size_t)g * c->Mgp + (size_t)c->h_cmds[s0 + k].row) * c->ws_ldx;
        }
    }
    hipEvent_t *ev = c->gev[g];
    auto &t = c->timing;
    const double upd_ms = elapsed(ev[0], ev[1]), dot_ms = elapsed(ev[4], ev[5]);
    t.lbfgs += upd_ms + dot_ms;
    t.update += upd_ms;
    t.dots += dot_ms;
    t.energy += elapsed(ev[1], ev[2]);
    t.softmax += elapsed(ev[2], ev[3]);
    t.grad += elapsed(ev[3], ev[4]);
    t.evals += 1;
    t.energy_launches += 1;
    t.grad_launches += 1;
    const double pb = (double)pair_bytes(c);
    double bu = 0.0, bd = 0.0;  // bytes of the update and dot-product passes
    double flops = 0.0;         // of one contraction
    for (int i = 0; i < c->gslots; ++i) {
        const sdca::SlotCmd &k = cmds[i];
        const int md = k.mode;
        if (md == sdca::MODE_IDLE) continue;
        const sdca_ctx *A = c->members[k.aln];
        const double nel = (double)A->q * (double)A->Kxp;  // elements of the command's vectors
        flops += 2.0 * A->B * (double)A->Kx * A->q;
        double u = 0.0, d = 0.0;  // bytes per element
        int npairs = 0;
        for (int p = 0; p < m; ++p) npairs += (k.live >> p) & 1u;
        if (md == sdca::MODE_INIT) u += 4.0 * (2 + (k.init_src != 0.f));
        else if (md == sdca::MODE_RETRY) u += 8.0 + pb;  // x_c and d read, x_t written
        else {
            // x, d and g_k read; x_t, d (and x_c) written; the new pair from g_old
            u += 4.0 * (3 + (k.accept != 0)) + 2.0 * pb + (k.pnew >= 0 ? 4.0 + 2.0 * pb : 0.0);
            for (int p = 0; p < m; ++p) {
                if (p == k.pnew) continue;
                u += (k.cs[p] != 0.f ? pb : 0.0) + (k.cy[p] != 0.f ? pb : 0.0);
            }
        }
        if (split) u += 2.0 * (k.level + 1);  // BF16 parts of x_t: hi (+ lo (+ l3))
        d += (md == sdca::MODE_INIT ? 4.0 : 8.0 + pb) + 2.0 * pb * npairs;  // g_t (, g_k, d) and the pairs
        bu += u * nel;
        bd += d * nel;
    }
    t.bytes_update += bu;
    t.bytes_dots += bd;
    t.bytes_lbfgs += bu + bd;
    t.flops_energy += flops;
    t.flops_grad += flops;
    return SDCA_OK;
}

// Reference layout (x of plmDCA_optimizer_parameters.hpp:176,219-247) -> device slot
// vector (q x Kxp floats).  Returns the FP64 sum of squares of the entries whose state
// is never observed (their only contribution is the L2 term).
double ref_to_dev(const sdca_ctx *c, int r, const double *x, float *dev)
{
    std::fill(dev, dev + c->n_slot, 0.f);
    double unobs = 0.0;
    const int L = c->L, q = c->q;
    for (int j = 0; j < L; ++j) {
        if (j == r) continue;
        for (int a = 0; a < q; ++a) {
            const int col = c->state_col[(size_t)j * q + a];
            const size_t o = sdca::ref_index(L, q, j, a, 0);
            for (int s = 0; s < q; ++s) {
                const double v = x[o + s];
                if (col >= 0) dev[(size_t)s * c->Kxp + col] = (float)v;
                else unobs += v * v;
            }
        }
    }
    for (int s = 0; s < q; ++s) dev[(size_t)s * c->Kxp + c->Kx] = (float)x[(size_t)L * q * q + s];
    return unobs;
}

void dev_to_ref(const sdca_ctx *c, int r, const float *dev, double *x)
{
    const int L = c->L, q = c->q;
    std::fill(x, x + (size_t)L * q * q + q, 0.0);
    for (int j = 0; j < L; ++j) {
        if (j == r) continue;
        for (int a = 0; a < q; ++a) {
            const int col = c->state_col[(size_t)j * q + a];
            if (col < 0) continue;
            const size_t o = sdca::ref_index(L, q, j, a, 0);
            for (int s = 0; s < q; ++s) x[o + s] = dev[(size_t)s * c->Kxp + col];
        }
    }
    for (int s = 0; s < q; ++s) x[(size_t)L * q * q + s] = dev[(size_t)s * c->Kxp + c->Kx];
}

sdca::SlotCmd idle_cmd()
{
    sdca::SlotCmd k;
    std::memset(&k, 0, sizeof(k));
    k.mode = sdca::MODE_IDLE;
    k.locus = -1;
    k.level = sdca::LV_SPLIT;
    k.item = -1;
    return k;
}

// ------------------------------------------------------- per-slot solver state

// PH_REBASE: the iterate is evaluated again at the next precision level.
enum Phase { PH_IDLE, PH_INIT, PH_LS, PH_REBASE, PH_DONE };

// Host side of one slot: line-search state and the inner products of the stored vectors
// that the two-loop recursion needs.  Pairs p hold s_p = alpha d and y_p = g_new - g_old;
// with g the current gradient, d the current direction and t the trial gradient:
//   SY[p][q] = s_p.y_q, YY[p][q] = y_p.y_q, SG[p] = s_p.g, YG[p] = y_p.g, GG = g.g,
//   DG = d.g, DY[p] = d.y_p (from d's coefficients and SY, YY, YG), and the trial's
//   TS[p], TY[p], TT, TG = t.g, TD = t.d.
struct Slot {
    int item = -1;  // index into the caller's loci list
    int aln = 0;    // alignment of the locus (member of the solving context)
    int r = -1;
    Phase phase = PH_IDLE;
    std::deque<int> pairs;  // pair slots, oldest first
    int gk = -1, gt = 0;    // gradient buffers: current (-1 before the first) and trial
    double f = 0.0;         // objective at the accepted iterate
    double stp = 0.0;       // current trial step
    double stp0 = 0.0;      // initial step of the current search
    double dg0 = 0.0;       // directional derivative at the start of the search
    sdca::MoreThuente ms;
    int ls_evals = 0;
    int nfeval = 0, iter = 0, status = -1;
    double gnorm = 0.0;
    bool from_xt = true;
    bool preloaded = false;  // the direction buffer holds a first search direction
    int level = sdca::LV_SPLIT;  // precision level of the slot's evaluations (sdca::Level)
    int restarts = 0;        // line searches restarted from steepest descent
    int nfeval_first = 0;    // evaluations up to the first accepted step
    int nfeval_fine = 0;     // evaluations at the fine level
    double gnormJ = 0.0;     // ||g||_inf over the coupling entries at the accepted iterate
    double SY[sdca::kMaxHistory][sdca::kMaxHistory], YY[sdca::kMaxHistory][sdca::kMaxHistory];
    double SG[sdca::kMaxHistory], YG[sdca::kMaxHistory], DY[sdca::kMaxHistory];
    double TS[sdca::kMaxHistory], TY[sdca::kMaxHistory];
    double GG = 0, DG = 0, TT = 0, TG = 0, TD = 0;
};

// A failed line search with curvature history is retried once more from steepest
// descent at the same iterate (at most this many times per locus) before the locus
// stops with SDCA_ST_LS_FAIL: near the FP32 rounding floor a stale history can give a
// direction along which no decrease is measurable while the gradient is not yet small.
constexpr int kMaxRestarts = 2;

// Steepest-descent steps (no curvature pairs yet) start at kFirstStep / ||g||_2.  The
// data term's curvature along g makes the accepted first step about 1/16 of 1/||g||_2 on
// the synthetic genomes: starting there takes the first line search from four
// evaluations to two (C4: median 62 instead of 63 evaluations per locus).
constexpr double kFirstStep = 1.0 / 16.0;

// Stopping rule of a Solver other than the context's options (the subspace warm start).
struct SolverLimits {
    double grad_threshold, j_threshold;
    int max_iter;
};

class Solver {
public:
    // gram: FP64 evaluations of the Gram formulation (one precision level, FP64 rounding of f)
    Solver(sdca_ctx *c, int nslots, int m, bool gram, const SolverLimits *lim = nullptr)
        : c_(c), n_(nslots), m_(m), slots_(nslots), cmds_(nslots, idle_cmd()),
          split_(!gram && c->opts.precision == SDCA_PREC_SPLIT_BF16), gram_(gram),
          gthr_(lim ? lim->grad_threshold : c->opts.grad_threshold),
          max_iter_(lim ? lim->max_iter : c->opts.max_iter),
          coarse_gnorm_(split_ && c->opts.coarse_factor > 0 ? c->opts.coarse_factor * c->opts.grad_threshold : 0.0),
          jthr_(lim ? lim->j_threshold : c->opts.j_threshold),
          fine_gnorm_(split_ && jthr_ > 0.0 ? c->opts.fine_factor * jthr_ : 0.0),
          max_level_(jthr_ > 0.0 && (!split_ || c->opts.fine_factor > 0) ? sdca::LV_FINE : sdca::LV_SPLIT)
    {
    }

    // Level a slot at `level` moves to once its ||g||_inf is ginf: coarse until
    // coarse_gnorm_, then split, then (with a coupling threshold) fine below fine_gnorm_.
    // FP32 contractions have no coarse or split level.
    int target_level(int level, double ginf) const
    {
        if (!split_) return max_level_;
        if (level == sdca::LV_COARSE && !(ginf >= coarse_gnorm_)) level = sdca::LV_SPLIT;
        if (level == sdca::LV_SPLIT && max_level_ == sdca::LV_FINE && ginf < fine_gnorm_) level = sdca::LV_FINE;
        return level;
    }

    // SuperDCA's rule (plmDCA.hpp:218-221) at full precision, plus the coupling threshold
    // at the highest level when one is set.
    bool converged(const Slot &S, double ginf, double gJ) const
    {
        if (!(ginf < gthr_) || S.level < sdca::LV_SPLIT) return false;
        return jthr_ <= 0.0 || (S.level == max_level_ && gJ < jthr_);
    }

    void assign(int i, int item, int aln, int r, bool warm, bool preloaded = false)
    {
        Slot &S = slots_[i];
        S = Slot();
        S.aln = aln;
        S.preloaded = preloaded;
        S.level = !split_ ? max_level_ : (coarse_gnorm_ > 0.0 ? sdca::LV_COARSE : sdca::LV_SPLIT);
        S.item = item;
        S.r = r;
        S.phase = PH_INIT;
        S.gt = 0;
        sdca::SlotCmd &k = cmds_[i];
        k = idle_cmd();
        k.mode = sdca::MODE_INIT;
        k.locus = r;
        k.gt = S.gt;
        k.gk = -1;
        k.pnew = -1;
        // the masked columns: the locus' own, or in a two-alignment scan every column of
        // the first alignment (plmDCA_objective.hpp:67: no block of the second is excluded)
        const sdca_ctx *A = c_->members[aln];
        k.aln = aln;
        k.self_lo = A->scan_split > 0 ? 0 : A->col_off[r];
        k.self_hi = A->scan_split > 0 ? A->col_off[A->scan_split] : A->col_off[r + 1];
        k.init_src = warm ? 1.f : 0.f;
        k.item = item;
    }

    void set_idle(int i)
    {
        slots_[i] = Slot();
        cmds_[i] = idle_cmd();
    }

    // Consumes the result row of slot i after a round.  Returns true when the slot's
    // locus is finished (status set).
    bool process(int i, const double *res)
    {
        Slot &S = slots_[i];
        sdca::SlotCmd &k = cmds_[i];
        const sdca::ResLayout R{m_};
        const int pending = k.mode;
        if (pending == sdca::MODE_NEWDIR) S.DG = res[R.dg()];  // d . g_k of the stored new direction
        for (int p : S.pairs) {
            S.TS[p] = res[R.ts(p)];
            S.TY[p] = res[R.ty(p)];
        }
        S.TT = res[R.tt()];
        S.TG = res[R.tg()];
        S.TD = res[R.td()];
        const sdca_ctx *A = c_->members[S.aln];
        const double f = res[R.fdata()] + A->lambda_J * res[R.xsqJ()] + A->lambda_h * res[R.xsqh()];
        const double ginf = res[R.ginf()], gJ = res[R.gJinf()];
        S.nfeval++;
        if (k.level == sdca::LV_FINE) S.nfeval_fine++;
        if (!std::isfinite(f) || !std::isfinite(ginf)) return finish(i, SDCA_ST_NONFINITE, S.phase == PH_INIT);
        if (S.phase == PH_INIT && target_level(S.level, ginf) > S.level) {
            // near the end already: evaluate the start point again at the higher level
            S.level = target_level(S.level, ginf);
            return false;
        }
        if (S.phase == PH_REBASE) {
            // the gradient at x_c at the new level becomes the current one
            std::swap(S.gk, S.gt);
            S.GG = S.TT;
            for (int p : S.pairs) {
                S.SG[p] = S.TS[p];
                S.YG[p] = S.TY[p];
            }
            S.f = f;
            S.gnorm = ginf;
            S.gnormJ = gJ;
            if (converged(S, ginf, gJ)) return finish(i, SDCA_ST_CONVERGED, true);  // x_t = x_c
            if (target_level(S.level, ginf) > S.level) return rebase(i);
            newdir(i, false, -1, 0.0);
            return false;
        }
        if (S.phase == PH_INIT) {
            S.gk = S.gt;
            S.gt = 1 - S.gk;
            S.GG = S.TT;
            S.f = f;
            S.gnorm = ginf;
            S.gnormJ = gJ;
            if (converged(S, ginf, gJ)) return finish(i, SDCA_ST_CONVERGED, true);
            if (S.preloaded) {
                S.preloaded = false;
                newdir_preloaded(i);
            } else {
                newdir(i, false, -1, 0.0);  // x_c already holds the start point
            }
            return false;
        }
        // line search in progress
        if (pending == sdca::MODE_NEWDIR) {
            const double dg0 = S.DG;
            if (!(dg0 < 0.0)) {  // rounding destroyed descent: restart from steepest descent
                S.pairs.clear();
                if (S.level < max_level_) return rebase(i);
                newdir(i, false, -1, 0.0);  // the trial point along d is not accepted
                return false;
            }
            S.dg0 = dg0;
            S.ms.start(S.f, dg0, S.stp0);
        }
        const double dg = S.TD;
        const double stp_trial = S.stp;
        // Approximate Wolfe conditions (W. W. Hager and H. Zhang, SIAM J. Optim. 16 (2005)
        // 170-192): near the optimum the objective, summed from FP32 energies, changes by
        // less than its rounding noise while the directional derivative is still accurate;
        // a step is then accepted on the derivative alone, with f allowed to exceed f(0)
        // by at most the noise level of the slot's precision.
        const double f_noise = gram_ ? 1e-13 : (split_ && S.level < sdca::LV_FINE ? 1e-5 : 1e-6);
        const bool approx_wolfe = dg >= S.ms.gtol * S.dg0 && dg <= (2.0 * S.ms.ftol - 1.0) * S.dg0 &&
                                  f <= S.f + f_noise * std::fabs(S.f);
        // Within the noise level the measured f no longer orders the trial steps: the search
        // then runs on phi(0) + stp (phi'(0) + phi'(stp)) / 2, the trapezoid rule over the
        // accurate directional derivatives (exact for a quadratic), so rounding in f cannot
        // fake a bracket.
        const double f_ls = std::fabs(f - S.f) <= f_noise * std::fabs(S.f) ? S.f + 0.5 * stp_trial * (S.dg0 + dg) : f;
        sdca::MoreThuente::Result act = sdca::MoreThuente::Converged;
        if (!approx_wolfe) {
            act = S.ms.step(f_ls, dg, S.stp);
            S.ls_evals++;
            if (act == sdca::MoreThuente::Continue && S.ls_evals < c_->opts.max_ls) {
                // a quasi-Newton step rejected twice in the coarse phase: the steps have
                // become small against the BF16 rounding of x, where the coarse objective
                // is too rough for the line search
                if (S.level == sdca::LV_COARSE && !S.pairs.empty() && S.ls_evals >= 2) return rebase(i);
                k.mode = sdca::MODE_RETRY;
                k.gt = S.gt;
                k.gk = S.gk;
                k.alpha = (float)S.stp;
                return false;
            }
            if (!(f_ls < S.f)) {
                // continue at the next level from x_c: the search is below this level's noise
                if (S.level < max_level_) return rebase(i);
                if (S.pairs.empty() || S.restarts >= kMaxRestarts) return finish(i, SDCA_ST_LS_FAIL, false);
                S.restarts++;
                S.pairs.clear();
                newdir(i, false, -1, 0.0);
                return false;
            }
        }
        // accept the trial: the new pair s = stp d, y = t - g (kept if s.y > 0)
        const double ys = stp_trial * (S.TD - S.DG);
        int pnew = -1;
        if (ys > 0.0) {
            if ((int)S.pairs.size() >= m_) {
                pnew = S.pairs.front();
                S.pairs.pop_front();
            } else {
                pnew = free_pair(i);
            }
            for (int q : S.pairs) {
                S.SY[pnew][q] = stp_trial * S.DY[q];
                S.SY[q][pnew] = S.TS[q] - S.SG[q];
                S.YY[pnew][q] = S.YY[q][pnew] = S.TY[q] - S.YG[q];
            }
            S.SY[pnew][pnew] = ys;
            S.YY[pnew][pnew] = S.TT - 2.0 * S.TG + S.GG;
            // products with the new current gradient t
            for (int q : S.pairs) {
                S.SG[q] = S.TS[q];
                S.YG[q] = S.TY[q];
            }
            S.SG[pnew] = stp_trial * S.TD;
            S.YG[pnew] = S.TT - S.TG;
            S.pairs.push_back(pnew);
        } else {
            S.pairs.clear();
        }
        S.GG = S.TT;
        std::swap(S.gk, S.gt);
        S.f = f;
        S.gnorm = ginf;
        S.gnormJ = gJ;
        S.iter++;
        if (S.iter == 1) S.nfeval_first = S.nfeval;
        const int lv = target_level(S.level, ginf);
        if (lv > S.level) {
            // end of a precision level: make the trial point the iterate (storing its pair)
            // and evaluate it again at the next level
            S.level = lv;
            newdir_identity(i, pnew, stp_trial);
            S.phase = PH_REBASE;
            return false;
        }
        if (converged(S, ginf, gJ)) return finish(i, SDCA_ST_CONVERGED, true);
        if (S.iter >= max_iter_) return finish(i, SDCA_ST_ITER_CAP, true);
        newdir(i, true, pnew, stp_trial);
        return false;
    }

    // End of a precision level without an accepted step: x_c is evaluated again at the
    // next level (a trial step of zero along the current direction).
    bool rebase(int i)
    {
        Slot &S = slots_[i];
        S.level = std::min<int>(max_level_, std::max(S.level + 1, target_level(S.level, S.gnorm)));
        S.phase = PH_REBASE;
        sdca::SlotCmd &k = cmds_[i];
        k.mode = sdca::MODE_RETRY;
        k.gt = S.gt;
        k.gk = S.gk;
        k.alpha = 0.f;
        return false;
    }

    // Marks the stored pairs of slots [i0, i1) for the coming round's products.
    void mark_live(int i0, int i1)
    {
        for (int i = i0; i < i1; ++i) {
            const Slot &S = slots_[i];
            sdca::SlotCmd &k = cmds_[i];
            k.live = 0u;
            k.level = S.level;
            if (S.phase == PH_IDLE || k.mode == sdca::MODE_IDLE) continue;
            for (int p : S.pairs) k.live |= 1u << p;
        }
    }

    const Slot &slot(int i) const { return slots_[i]; }
    const std::vector<sdca::SlotCmd> &cmds() const { return cmds_; }
    std::vector<sdca::SlotCmd> &cmds() { return cmds_; }

private:
    bool finish(int i, int status, bool from_xt)
    {
        Slot &S = slots_[i];
        S.status = status;
        S.phase = PH_DONE;
        S.from_xt = from_xt;
        return true;
    }

    int free_pair(int i) const
    {
        const Slot &S = slots_[i];
        for (int p = 0; p < m_; ++p) {
            bool used = false;
            for (int q : S.pairs) used |= q == p;
            if (!used) return p;
        }
        return -1;
    }

    // d = -H g as coefficients over {g, s_p, y_p} (lbfgs_coeffs.hpp)
    void two_loop(int i, double *cs, double *cy, double &cg)
    {
        const Slot &S = slots_[i];
        int pairs[sdca::kMaxHistory];
        const int n = (int)S.pairs.size();
        for (int t = 0; t < n; ++t) pairs[t] = S.pairs[t];
        sdca::two_loop_coeffs(pairs, n, m_, S.SY, S.YY, S.SG, S.YG, cs, cy, cg);
    }

    // First direction = the preloaded direction buffer, unit initial step.  If it turns
    // out not to be a descent direction the line search start falls back to steepest
    // descent.
    void newdir_preloaded(int i)
    {
        Slot &S = slots_[i];
        S.stp0 = 1.0;
        S.stp = 1.0;
        S.ls_evals = 0;
        S.phase = PH_LS;
        sdca::SlotCmd &k = cmds_[i];
        k.mode = sdca::MODE_NEWDIR;
        k.gt = S.gt;
        k.gk = S.gk;
        k.accept = 0;  // x_c already holds the start point
        k.pnew = -1;
        k.pstep = 0.f;
        k.alpha = 1.f;
        k.cd = 1.f;
        k.cg = 0.f;
        for (int p = 0; p < sdca::kMaxHistory; ++p) k.cs[p] = k.cy[p] = 0.f;
    }

    // x_c <- x_c + pstep d (the accepted trial point) with its pair pnew (if any), the
    // direction kept, and a trial step of zero: the round evaluates the new iterate.
    void newdir_identity(int i, int pnew, double pstep)
    {
        Slot &S = slots_[i];
        if (pnew >= 0) S.DY[pnew] = S.TD - S.DG;  // d . (g_new - g_old); the older d . y_p stand
        sdca::SlotCmd &k = cmds_[i];
        k.mode = sdca::MODE_NEWDIR;
        k.gt = slots_[i].gt;
        k.gk = slots_[i].gk;
        k.accept = 1;
        k.pnew = pnew;
        k.pstep = (float)pstep;
        k.alpha = 0.f;
        k.cd = 1.f;
        k.cg = 0.f;
        for (int p = 0; p < sdca::kMaxHistory; ++p) k.cs[p] = k.cy[p] = 0.f;
    }

    // New search direction from the iterate: the last trial point x_c + pstep d if accept
    // (with the new pair pnew of that step, if any), else the current x_c (the start
    // point, or a restart after a failed or non-descent search).
    void newdir(int i, bool accept, int pnew, double pstep)
    {
        Slot &S = slots_[i];
        double cs[sdca::kMaxHistory], cy[sdca::kMaxHistory], cg = -1.0;
        bool ok = false;
        if (!S.pairs.empty()) {
            two_loop(i, cs, cy, cg);
            double dg = cg * S.GG;
            bool finite = std::isfinite(cg);
            for (int p : S.pairs) {
                dg += cs[p] * S.SG[p] + cy[p] * S.YG[p];
                finite &= std::isfinite(cs[p]) && std::isfinite(cy[p]);
            }
            ok = finite && dg < 0.0;
            if (!ok) S.pairs.clear();
        }
        if (!ok) {
            for (int p = 0; p < m_; ++p) cs[p] = cy[p] = 0.0;
            cg = -1.0;
        }
        S.stp0 = S.pairs.empty() ? kFirstStep / std::sqrt(S.GG) : 1.0;
        S.stp = S.stp0;
        S.ls_evals = 0;
        S.phase = PH_LS;
        sdca::SlotCmd &k = cmds_[i];
        k.mode = sdca::MODE_NEWDIR;
        k.gt = S.gt;
        k.gk = S.gk;
        k.accept = accept ? 1 : 0;
        k.pnew = pnew;
        k.pstep = (float)pstep;
        k.alpha = (float)S.stp;
        k.cd = 0.f;
        k.cg = (float)cg;
        for (int p = 0; p < sdca::kMaxHistory; ++p) {
            k.cs[p] = p < m_ ? (float)cs[p] : 0.f;
            k.cy[p] = p < m_ ? (float)cy[p] : 0.f;
        }
        // d . y_p of the new direction from the pair products, with the coefficients as
        // the device applies them: d = cg g_k + sum_q (cs_q s_q + cy_q y_q)
        for (int p : S.pairs) {
            double v = (double)k.cg * S.YG[p];
            for (int q : S.pairs) v += (double)k.cs[q] * S.SY[q][p] + (double)k.cy[q] * S.YY[q][p];
            S.DY[p] = v;
        }
    }

    sdca_ctx *c_;
    int n_, m_;
    std::vector<Slot> slots_;
    std::vector<sdca::SlotCmd> cmds_;
    bool split_;  // split-BF16 contractions (else FP32)
    bool gram_;   // Gram formulation: FP64 evaluations
    double gthr_;   // ||g|| threshold of the stopping rule
    int max_iter_;  // iteration cap per locus
    // split precision: coarse phase (BF16 high parts only) while ||g||_inf >= this; 0: none
    double coarse_gnorm_;
    double jthr_;        // coupling-gradient threshold (0: none)
    double fine_gnorm_;  // split precision with jthr_: the fine level starts below this ||g||_inf
    int max_level_;      // LV_FINE with a coupling threshold, else LV_SPLIT
};

// ------------------------------------------------------- host worker threads

// A fixed set of worker threads for the host's per-slot L-BFGS steps between rounds:
// run(lo, hi, fn) calls fn(i) for every i in [lo, hi), split into one contiguous range per
// thread (the calling thread takes the first).  Each slot's step touches that slot's state
// and command alone, so the ranges are independent; the result is the same as a serial loop.
class SlotPool {
public:
    explicit SlotPool(int nthreads) : nt_(std::max(1, nthreads))
    {
        for (int t = 1; t < nt_; ++t) th_.emplace_back([this, t] { worker(t); });
    }
    ~SlotPool()
    {
        {
            std::lock_guard<std::mutex> l(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    void run(int lo, int hi, const std::function<void(int)> &fn)
    {
        if (nt_ <= 1 || hi - lo < 256) {
            for (int i = lo; i < hi; ++i) fn(i);
            return;
        }
        {
            std::lock_guard<std::mutex> l(mu_);
            fn_ = &fn;
            lo_ = lo;
            hi_ = hi;
            pending_ = nt_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        range(0, lo, hi, fn);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return pending_ == 0; });
        fn_ = nullptr;
    }
    // the worker count for this process: its CPUs (the affinity mask), at most 8
    static int default_threads()
    {
        cpu_set_t set;
        int n = 1;
        if (sched_getaffinity(0, sizeof(set), &set) == 0) n = CPU_COUNT(&set);
        return std::max(1, std::min(8, n));
    }

private:
    void range(int t, int lo, int hi, const std::function<void(int)> &fn) const
    {
        const int64_t n = hi - lo;
        const int a = lo + (int)(n * t / nt_), b = lo + (int)(n * (t + 1) / nt_);
        for (int i = a; i < b; ++i) fn(i);
    }
    void worker(int t)
    {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)> *fn;
            int lo, hi;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                fn = fn_;
                lo = lo_;
                hi = hi_;
            }
            range(t, lo, hi, *fn);
            {
                std::lock_guard<std::mutex> l(mu_);
                if (--pending_ == 0) done_.notify_one();
            }
        }
    }
    int nt_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)> *fn_ = nullptr;
    int lo_ = 0, hi_ = 0, pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

// Host threads of a Gram solve: with four slot groups, two threads step two groups each
// (one host thread does not step four groups as fast as the device runs their rounds);
// otherwise the calling thread steps them all.  The first error stops the other thread.
class HostThreads {
public:
    HostThreads(int ngroups, int device) : ng_(ngroups), nt_(ngroups == 4 ? 2 : 1), device_(device) {}

    // fn(g0, g1, msg) steps groups [g0, g1) and returns an sdca status (msg: its error text).
    void run(const std::function<int(int, int, std::string &)> &fn)
    {
        const int per = ng_ / nt_;
        std::vector<int> rc(nt_, SDCA_OK);
        std::vector<std::string> msg(nt_);
        auto body = [&](int t) {
            if (t > 0 && hipSetDevice(device_) != hipSuccess) {
                rc[t] = SDCA_E_HIP;
                msg[t] = "hipSetDevice failed on a host thread";
            } else {
                rc[t] = fn(t * per, (t + 1) * per, msg[t]);
                if (rc[t] != SDCA_OK && msg[t].empty()) msg[t] = g_err;  // this thread's error text
            }
            if (rc[t] != SDCA_OK) stop_ = true;
        };
        std::vector<std::thread> th;
        for (int t = 1; t < nt_; ++t) th.emplace_back(body, t);
        body(0);
        for (std::thread &x : th) x.join();
        for (int t = 0; t < nt_; ++t)
            if (rc[t] != SDCA_OK) {
                rc_ = rc[t];
                msg_ = msg[t];
                break;
            }
    }
    bool stopped() const { return stop_.load(); }
    int result() const { return rc_; }
    const std::string &message() const { return msg_; }

private:
    int ng_, nt_, device_;
    std::atomic<bool> stop_{false};
    int rc_ = SDCA_OK;
    std::string msg_;
};

// ------------------------------------------------------------ Gram formulation
//
// The solve in the representation of sdca_gram.hip: FP64 slot vectors of coefficients,
// their images under K_r and the fields; per round one trial-point/softmax pass, one
// B x B x q GEMM per alignment, one gradient-image/products pass, and the same host Solver
// as the primal formulation.  A finished locus' couplings J = X_r^T c are formed once, by
// the gradient contraction of the primal path without its L2 term, and scored.

bool gram_wanted(const sdca_ctx *c, bool warm)
{
    const int f = c->opts.formulation;
    if (f == SDCA_FORM_PRIMAL || warm) return false;
    if (f == SDCA_FORM_GRAM) return true;
    for (const sdca_ctx *A : c->members)
        if ((int64_t)A->B > A->Kx) return false;
    return true;
}

// Identical-locus counts of alignment A over the loci its energies use: all of them (a
// command removes its own locus), or in a two-alignment scan those of the second alignment
// (plmDCA_objective.hpp:67: the first alignment's blocks are excluded).
int ensure_gram_matrix(sdca_ctx *A, hipStream_t st)
{
    const int j0 = A->scan_split > 0 ? A->scan_split : 0;
    if (A->d_gram && A->gram_split == j0) return SDCA_OK;
    dfree(A->d_gram);
    A->gram_split = -1;
    A->sub_rank = 0;  // the subspace basis belongs to the old counts
    int rc;
    if ((rc = dalloc(&A->d_gram, (size_t)A->Bp * A->Bp))) return rc;
    const int W = (A->L + 63) / 64;
    uint64_t *planes = nullptr;
    if ((rc = dalloc(&planes, (size_t)A->B * A->q * W))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    sdca::kern::bitplanes(st, A->d_states, A->B, A->L, A->q, W, planes);
    sdca::gram::gram_matrix(st, planes, A->B, A->q, W, j0, A->Bp, A->d_gram);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dfree(planes);
    if (e != hipSuccess) return fail(SDCA_E_HIP, "identical-locus counts failed: %s", hipGetErrorString(e));
    A->gram_split = j0;
    return SDCA_OK;
}

int ensure_gram_ws(sdca_ctx *c, int nslots, int m)
{
    int gq = 0, gBp = 0, gL = 0;
    int64_t nitem = 0;
    for (const sdca_ctx *A : c->members) {
        gq = std::max(gq, A->q);
        gBp = std::max(gBp, A->Bp);
        gL = std::max(gL, A->L);
        nitem = std::max(nitem, (int64_t)A->q * A->Kxp);
    }
    const int nmem = (int)c->members.size();
    // two slot groups (their rounds overlap: one group's GEMM with the other's vector passes
    // and host steps) unless pipeline_groups is 1 or the batch is small
    const int pg = c->opts.pipeline_groups;
    const int ng = (pg == 1 || nslots < 64) ? 1 : (pg == 4 && nslots >= 256 ? 4 : 2);
    sdca::GramWs *G = c->gram;
    if (G && G->slots == nslots && G->m == m && G->gq == gq && G->gBp == gBp && G->gL == gL && G->nmem == nmem &&
        G->n_item == nitem && G->ngroups == ng)
        return SDCA_OK;
    delete c->gram;
    c->gram = G = new sdca::GramWs();
    G->slots = nslots;
    G->m = m;
    G->gq = gq;
    G->gBp = gBp;
    G->gL = gL;
    G->nmem = nmem;
    G->n_item = nitem;
    G->nvs = sdca::gram::vec_len(gq, gBp);
    G->vstride = (sdca::gram::kVecPairs + 2 * m) * G->nvs;
    G->ldG = gBp;
    G->nchunk = gBp / 256;
    G->nchq = sdca::gram::rows(gq) * gBp / 256;  // blocks of 256 over the coordinates of a vector half
    G->nchd = sdca::gram::dots_chunks((int64_t)sdca::gram::rows(gq) * gBp);
    G->nmat = std::min(nslots, 128);
    G->ngroups = ng;
    G->gs = (nslots + ng - 1) / ng;
    // the GEMM reads whole tiles of kGemmCols columns
    G->colcap = round_up((int64_t)G->gs * gq, sdca::gram::kGemmCols) + sdca::gram::kGemmCols;
    // each alignment's rows of a batch of finished loci start on a tile boundary
    G->mrows = (int)round_up((int64_t)G->nmat * gq, sdca::kTile) + std::min(nmem, G->nmat) * sdca::kTile;
    const int W = sdca::ResLayout{m}.width();
    const int64_t ncols = G->colcap * ng;
    const int npad = G->gs * ng;  // per-command buffers, a group's from its first slot on
    int rc;
    if ((rc = dalloc(&G->V, (size_t)nslots * G->vstride)) || (rc = dalloc(&G->Gp, (size_t)ncols * G->ldG)) ||
        (rc = dalloc(&G->Zp, (size_t)ncols * G->ldG)) ||
        (rc = dalloc(&G->upd_part, (size_t)npad * (G->nchq + 1) * 3)) ||
        (rc = dalloc(&G->sm_part, (size_t)npad * G->nchunk * 65)) ||
        (rc = dalloc(&G->dot_part, (size_t)npad * G->nchd * sdca::ResLayout{m}.ndots())) ||
        (rc = dalloc(&G->res, (size_t)npad * W)) || (rc = dalloc(&G->d_cmds, (size_t)npad)) ||
        (rc = dalloc(&G->d_galn, (size_t)nmem)) || (rc = dalloc(&G->d_aln, (size_t)nmem)))
        return rc;
    // the coupling contraction of finished loci: two BF16 parts (hi + lo, 16 bits of each
    // coordinate) for the scores, three (hi + mid + lo) for x_out, at either precision setting
    const size_t mel = (size_t)G->mrows * gBp;
    if ((rc = dalloc(&G->Mhi, mel)) || (rc = dalloc(&G->Mlo, mel)) || (rc = dalloc(&G->Ml3, mel))) return rc;
    G->pcap = G->nmat + G->gs;
    if ((rc = dalloc(&G->Pend, (size_t)2 * G->pcap * G->nvs))) return rc;
    // score rows of one flush: a flush of more loci than this runs synchronously
    G->nhrows = std::min(G->pcap, 4 * G->nmat);
    if ((rc = dalloc(&G->Mout, (size_t)G->nmat * nitem)) || (rc = dalloc(&G->mhpart, (size_t)G->nmat * 8)) ||
        (rc = dalloc(&G->d_mcmds, (size_t)G->nmat)) || (rc = dalloc(&G->d_mgroups, (size_t)nmem)) ||
        (rc = dalloc(&G->d_moff, (size_t)G->nmat)) || (rc = dalloc(&G->d_soff, (size_t)G->nmat)) ||
        (rc = dalloc(&G->d_mrow, (size_t)G->nmat)) || (rc = dalloc(&G->d_maln, (size_t)G->nmat)) ||
        (rc = dalloc(&G->d_sloc, (size_t)G->nmat)) || (rc = dalloc(&G->d_sld, (size_t)G->nmat)) ||
        (rc = dalloc(&G->d_rows, (size_t)G->nhrows * gL)))
        return rc;
    if ((rc = dalloc(&G->d_stage, (size_t)npad * 5))) return rc;
    HIP_TRY(hipHostMalloc((void **)&G->h_stage, (size_t)npad * 5 * sizeof(int64_t), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&G->h_cmds, (size_t)npad * sizeof(sdca::SlotCmd), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&G->h_res, (size_t)npad * W * sizeof(double), hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void **)&G->h_rows, (size_t)G->nhrows * gL * sizeof(float), hipHostMallocDefault));
    for (auto &g : G->ev)
        for (auto &e : g) HIP_TRY(hipEventCreate(&e));
    for (auto &e : G->stage_ev) HIP_TRY(hipEventCreate(&e));
    for (auto &e : G->pend_free) HIP_TRY(hipEventCreate(&e));
    for (auto &e : G->fev) HIP_TRY(hipEventCreate(&e));
    G->pend.clear();
    G->pend.reserve(G->pcap);
    HIP_TRY(hipDeviceSynchronize());
    return SDCA_OK;
}

// Per solve: the members' FP64 weights and descriptors (weights and lambda may change
// between solves).
int upload_gram_alns(sdca_ctx *c)
{
    sdca::GramWs *G = c->gram;
    std::vector<sdca::GramAln> v(c->members.size());
    int rc;
    for (size_t a = 0; a < c->members.size(); ++a) {
        sdca_ctx *A = c->members[a];
        if (!A->d_w64 && (rc = dalloc(&A->d_w64, (size_t)A->Bp))) return rc;
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy(A->d_w64, A->h_w.data(), (size_t)A->B * sizeof(double), hipMemcpyHostToDevice));
        sdca::GramAln &g = v[a];
        g.K = A->d_gram;
        g.statesT = A->d_statesT;
        g.w = A->d_w64;
        g.B = A->B;
        g.Bp = A->Bp;
        g.q = A->q;
        g.own = A->scan_split > 0 ? 0 : 1;
        g.lamJ2 = 2.0 * A->lambda_J;
        g.lamh2 = 2.0 * A->lambda_h;
    }
    HIP_TRY(hipMemcpy(G->d_galn, v.data(), v.size() * sizeof(sdca::GramAln), hipMemcpyHostToDevice));
    return upload_alns(c, G->d_aln);
}

// One round of slot group g (slots [g gs, (g + 1) gs)) on its stream: the commands of its
// active slots (by alignment, in slot order), each owning q consecutive columns of the
// group's share of the packed GEMM operand.
// With graph_rounds (one alignment): every slot of the group gets a command (idle ones
// stay IDLE), command k at columns k q; the round was captured once as a HIP graph with
// these fixed grids and is replayed.
int gram_launch_graph(sdca_ctx *c, int g, const std::vector<sdca::SlotCmd> &cmds)
{
    sdca::GramWs *G = c->gram;
    const int m = G->m, W = sdca::ResLayout{m}.width();
    const sdca_ctx *A = c->members[0];
    const int s0 = g * G->gs, s1 = std::min<int>((int)cmds.size(), s0 + G->gs), n = s1 - s0;
    std::vector<int> &lst = G->cmd_slot[g];
    lst.clear();
    sdca::SlotCmd *hc = G->h_cmds + s0;
    for (int i = s0; i < s1; ++i) {
        sdca::SlotCmd &k = hc[i - s0];
        k = cmds[i];
        k.slot = i;
        k.row = (i - s0) * sdca::gram::rows(A->q);
        k.aln = 0;
        k.q = A->q;
        k.Kx = (int)A->Kx;
        k.Kxp = (int)A->Kxp;
        lst.push_back(i);
    }
    hipStream_t st = c->gstream[g];
    hipEvent_t *ev = G->ev[g];
    if (!G->gexec[g] || G->graph_K[g] != A->d_gram) {
        if (G->gexec[g]) (void)hipGraphExecDestroy(G->gexec[g]);
        if (G->graph[g]) (void)hipGraphDestroy(G->graph[g]);
        G->gexec[g] = nullptr;
        G->graph[g] = nullptr;
        sdca::SlotCmd *dc = G->d_cmds + s0;
        double *Gp = G->Gp + (size_t)g * G->colcap * G->ldG, *Zp = G->Zp + (size_t)g * G->colcap * G->ldG;
        double *upd = G->upd_part + (size_t)s0 * (G->nchq + 1) * 3, *sm = G->sm_part + (size_t)s0 * G->nchunk * 65;
        double *dot = G->dot_part + (size_t)s0 * G->nchd * sdca::ResLayout{m}.ndots(), *res = G->res + (size_t)s0 * W;
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        (void)hipMemcpyAsync(dc, hc, n * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st);
        sdca::gram::step(st, m, dc, n, G->d_galn, G->V, G->vstride, G->nvs, Gp, G->ldG, upd, sm, G->nchunk, G->nchq,
                         nullptr, nullptr);
        sdca::gram::gemm(st, A->d_gram, A->Bp, Gp, G->ldG, Zp, G->ldG, A->Bp, n * sdca::gram::rows(A->q), A->Bp);
        sdca::gram::dots(st, m, dc, n, G->d_galn, G->V, G->vstride, G->nvs, Zp, G->ldG, sm, G->nchunk, dot, G->nchd,
                         0);
        sdca::gram::reduce(st, m, n, upd, dot, sm, G->nchunk, G->nchq, G->nchd, res);
        (void)hipMemcpyAsync(G->h_res + (size_t)s0 * W, res, (size_t)n * W * sizeof(double), hipMemcpyDeviceToHost,
                             st);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &graph);
        if (e != hipSuccess) return fail(SDCA_E_HIP, "capturing a round failed: %s", hipGetErrorString(e));
        G->graph[g] = graph;
        HIP_TRY(hipGraphInstantiate(&G->gexec[g], graph, nullptr, nullptr, 0));
        G->graph_K[g] = A->d_gram;
    }
    HIP_TRY(hipEventRecord(ev[0], st));
    HIP_TRY(hipGraphLaunch(G->gexec[g], st));
    HIP_TRY(hipEventRecord(ev[3], st));
    return SDCA_OK;
}

int gram_launch_round(sdca_ctx *c, int g, const std::vector<sdca::SlotCmd> &cmds)
{
    sdca::GramWs *G = c->gram;
    const int m = G->m, W = sdca::ResLayout{m}.width();
    const int nmem = (int)c->members.size();
    if (c->opts.graph_rounds && nmem == 1 && !G->warm_rounds) return gram_launch_graph(c, g, cmds);
    const int s0 = g * G->gs, s1 = std::min<int>((int)cmds.size(), s0 + G->gs);
    std::vector<int> &lst = G->cmd_slot[g];
    lst.clear();
    sdca::SlotCmd *hc = G->h_cmds + s0;
    std::vector<int> col0(nmem, 0), ncol(nmem, 0);
    int col = 0, nwarm = 0;
    auto staged = [&](const sdca::SlotCmd &k) {
        return G->warm_rounds && k.mode == sdca::MODE_INIT && k.init_src != 0.f && k.item >= 0;
    };
    for (int a = 0; a < nmem; ++a) {
        const sdca_ctx *A = c->members[a];
        col0[a] = col;
        // (warm rounds, one alignment) the staged starts first: the start kernel covers them alone
        for (int pass = 0; pass < 2; ++pass)
            for (int i = s0; i < s1; ++i) {
                if (cmds[i].mode == sdca::MODE_IDLE || cmds[i].aln != a || staged(cmds[i]) != (pass == 0)) continue;
                nwarm += pass == 0;
                sdca::SlotCmd &k = hc[lst.size()];
                k = cmds[i];
                k.slot = i;
                k.row = col;
                k.q = A->q;
                k.Kx = (int)A->Kx;
                k.Kxp = (int)A->Kxp;
                lst.push_back(i);
                col += sdca::gram::rows(A->q);
            }
        ncol[a] = col - col0[a];
    }
    const int nc = (int)lst.size();
    if (nc == 0) return SDCA_OK;
    hipStream_t st = c->gstream[g];
    hipEvent_t *ev = G->ev[g];
    sdca::SlotCmd *dc = G->d_cmds + s0;
    double *Gp = G->Gp + (size_t)g * G->colcap * G->ldG, *Zp = G->Zp + (size_t)g * G->colcap * G->ldG;
    double *upd = G->upd_part + (size_t)s0 * (G->nchq + 1) * 3, *sm = G->sm_part + (size_t)s0 * G->nchunk * 65;
    double *dot = G->dot_part + (size_t)s0 * G->nchd * sdca::ResLayout{m}.ndots(), *res = G->res + (size_t)s0 * W;
    HIP_TRY(hipMemcpyAsync(dc, hc, nc * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ev[0], st));
    if (nwarm > 0)  // staged starts of the round's new loci into their x_c
        sdca::sub::warm(st, dc, nwarm, G->d_galn, G->V, G->vstride, G->nvs, G->Cst, G->Kst, G->Wst + G->wfld);
    sdca::gram::step(st, m, dc, nc, G->d_galn, G->V, G->vstride, G->nvs, Gp, G->ldG, upd, sm, G->nchunk, G->nchq,
                     ev[6], ev[7]);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[1], st));
    for (int a = 0; a < nmem; ++a) {
        if (!ncol[a]) continue;
        const sdca_ctx *A = c->members[a];
        sdca::gram::gemm(st, A->d_gram, A->Bp, Gp + (size_t)col0[a] * G->ldG, G->ldG, Zp + (size_t)col0[a] * G->ldG,
                         G->ldG, A->Bp, ncol[a], A->Bp);
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[2], st));
    sdca::gram::dots(st, m, dc, nc, G->d_galn, G->V, G->vstride, G->nvs, Zp, G->ldG, sm, G->nchunk, dot, G->nchd,
                     0);
    sdca::gram::reduce(st, m, nc, upd, dot, sm, G->nchunk, G->nchq, G->nchd, res);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[3], st));
    HIP_TRY(hipMemcpyAsync(G->h_res + (size_t)s0 * W, res, (size_t)nc * W * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    return SDCA_OK;
}

// Waits for group g's round, moves its result rows to their slots (the norms: square roots
// of the squared 2-norms the device summed) and accounts its time, flops and bytes.
int gram_finish_round(sdca_ctx *c, int g, const std::vector<sdca::SlotCmd> &cmds, std::vector<double> &res)
{
    sdca::GramWs *G = c->gram;
    const std::vector<int> &lst = G->cmd_slot[g];
    if (lst.empty()) return SDCA_OK;
    HIP_TRY(hipStreamSynchronize(c->gstream[g]));
    const sdca::ResLayout R{G->m};
    const int W = R.width();
    const double *hr = G->h_res + (size_t)g * G->gs * W;
    for (size_t k = 0; k < lst.size(); ++k) {
        double *o = res.data() + (size_t)lst[k] * W;
        std::memcpy(o, hr + k * W, W * sizeof(double));
        o[R.ginf()] = std::sqrt(std::max(0.0, o[R.ginf()]));
        o[R.gJinf()] = std::sqrt(std::max(0.0, o[R.gJinf()]));
    }
    hipEvent_t *ev = G->ev[g];
    std::lock_guard<std::mutex> lk(G->mu);  // the timing is shared by the host threads
    auto &t = c->timing;
    const bool graphed = c->opts.graph_rounds && c->members.size() == 1 && !G->warm_rounds;
    if (graphed) {
        t.lbfgs += elapsed(ev[0], ev[3]);  // the whole replayed round
    } else {
        const double upd = elapsed(ev[0], ev[1]), gm = elapsed(ev[1], ev[2]), dts = elapsed(ev[2], ev[3]);
        t.gstep += elapsed(ev[6], ev[7]);
        t.update += upd;
        t.grad += gm;
        t.dots += dts;
        t.lbfgs += upd + dts;
        t.grad_launches += 1;
    }
    t.evals += 1;
    double flops = 0.0, bu = 0.0, bd = 0.0, bgs = 0.0;
    for (int i : lst) {
        const sdca::SlotCmd &k = cmds[i];
        if (k.mode == sdca::MODE_IDLE) continue;
        const sdca_ctx *A = c->members[k.aln];
        const double half = (double)sdca::gram::rows(A->q) * A->Bp;  // coordinates (or image entries) of a vector
        flops += 2.0 * A->B * (double)A->B * sdca::gram::rows(A->q);
        int npairs = 0;
        for (int p = 0; p < G->m; ++p) npairs += (k.live >> p) & 1u;
        // vector accesses (coefficients and image) of the trial-point pass, 8 bytes each
        double acc;
        if (k.mode == sdca::MODE_INIT) acc = 2;
        else if (k.mode == sdca::MODE_RETRY) acc = 3;
        else {
            acc = 3 + (k.accept != 0) + (k.pnew >= 0 ? 3 : 0) + 2;
            for (int p = 0; p < G->m; ++p) {
                if (p == k.pnew) continue;
                acc += (k.cs[p] != 0.f) + (k.cy[p] != 0.f);
            }
        }
        // + the softmax: x_t read, the beliefs and the gradient's coefficients written
        bu += 8.0 * half * (2.0 * acc + 4.0);
        // k_gstep alone: the accesses of both halves
        bgs += 8.0 * half * 2.0 * acc;
        // products pass: Z, the trial image and gradient coefficients read, its image written,
        // and the images of g_k, d and the live pairs
        bd += 8.0 * half * (4.0 + (k.mode == sdca::MODE_INIT ? 0.0 : 2.0 + 2.0 * npairs));
    }
    t.flops_grad += flops;
    t.bytes_update += bu;
    t.bytes_dots += bd;
    t.bytes_gstep += bgs;
    t.bytes_lbfgs += bu + bd;
    return SDCA_OK;
}

// Finished slots fin of group g: their statistics, and their solutions (x_t or x_c:
// coefficients and fields) copied to the pending buffer on the group's stream, so that the
// slots can take new loci at once.
int gram_stage(sdca_ctx *c, int g, const Solver &S, const std::vector<int> &fin, sdca_locus_stats *st_out)
{
    sdca::GramWs *G = c->gram;
    hipStream_t st = c->gstream[g];
    const int nf = (int)fin.size();
    int64_t *meta = G->h_stage + (size_t)g * G->gs * 5;
    int64_t maxlen = 0;
    for (int k = 0; k < nf; ++k) {
        const int i = fin[k];
        const Slot &sl = S.slot(i);
        const sdca_ctx *A = c->members[sl.aln];
        const int64_t qB = (int64_t)sdca::gram::rows(A->q) * A->Bp;
        const size_t u = G->pend.size();
        if ((int)u >= G->pcap) return fail(SDCA_E_STATE, "pending buffer full");
        // coefficients and fields of x_t or x_c into the pending buffer
        meta[k] = (int64_t)i * G->vstride + (sl.from_xt ? sdca::gram::kVecXt : sdca::gram::kVecXc) * G->nvs;
        meta[nf + k] = ((int64_t)G->phalf * G->pcap + (int64_t)u) * G->nvs;
        meta[2 * nf + k] = qB;
        meta[3 * nf + k] = 2 * qB;
        meta[4 * nf + k] = 2 * qB;
        maxlen = std::max(maxlen, qB);
        G->pend.push_back({sl.item, sl.r, sl.aln});
        if (st_out) {
            sdca_locus_stats &o = st_out[sl.item];
            o.fval = sl.f;
            o.gnorm = sl.gnorm;
            o.nfeval = sl.nfeval;
            o.iter = sl.iter;
            o.status = sl.status;
            o.nfeval_first = sl.nfeval_first;
            o.gnorm_J = sl.gnormJ;
            o.nfeval_fine = 0;
            o.nfeval_sub = G->sub_nev ? G->sub_nev[sl.item] : 0;
        }
    }
    if (nf > 0) {
        int64_t *dmeta = G->d_stage + (size_t)g * G->gs * 5;
        HIP_TRY(hipMemcpyAsync(dmeta, meta, (size_t)nf * 5 * sizeof(int64_t), hipMemcpyHostToDevice, st));
        HIP_TRY(hipStreamWaitEvent(st, G->pend_free[G->phalf], 0));  // the half's last flush has read it
        sdca::gram::gather(st, G->V, G->Pend, dmeta, nf, maxlen, 8);
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipEventRecord(G->stage_ev[g], st));
    return SDCA_OK;
}

// The score rows of the last flush into score_rows (or dropped, when score_rows is null),
// once its batches have run; the copy in pieces of rows on the pool's threads (about 48 MB
// per 128 loci at C4: one thread would hold the rounds up for milliseconds).
int gram_consume(sdca_ctx *c, float *score_rows, const int64_t *row_off, SlotPool *pool)
{
    sdca::GramWs *G = c->gram;
    if (!G->finflight) return SDCA_OK;
    G->finflight = false;
    HIP_TRY(hipStreamSynchronize(c->fstream));
    c->timing.score += elapsed(G->fev[0], G->fev[1]);
    if (score_rows) {
        constexpr int kPieces = 8;
        const std::vector<sdca::GramWs::RowOut> &F = G->fout;
        auto piece = [&](int u) {
            const sdca::GramWs::RowOut &o = F[u / kPieces];
            const int La = c->members[o.aln]->L, p = u % kPieces;
            const int64_t off = row_off ? row_off[o.item] : (int64_t)o.item * La;
            const int64_t a = (int64_t)La * p / kPieces, b = (int64_t)La * (p + 1) / kPieces;
            std::memcpy(score_rows + off + a, G->h_rows + (size_t)o.hrow * G->gL + a, (size_t)(b - a) * 4);
        };
        const int n = (int)F.size() * kPieces;
        if (pool) pool->run(0, n, piece);
        else
            for (int u = 0; u < n; ++u) piece(u);
    }
    G->fout.clear();
    return SDCA_OK;
}

// Couplings, score rows and x_out of the pending loci (one half of Pend), nmat per batch on
// the coupling stream after every group's staging copies; all of them once nmat have
// gathered, or when force is set.  The rounds go on meanwhile: the rows are copied out at
// the next flush or by gram_consume at the end of the solve.  With x_out, or more pending
// loci than the row buffers hold, each batch is waited for and copied out at once.
int gram_flush(sdca_ctx *c, bool force, float *score_rows, const int64_t *row_off, double *x_out, SlotPool *pool)
{
    sdca::GramWs *G = c->gram;
    if (G->pend.empty() || (!force && (int)G->pend.size() < G->nmat)) return SDCA_OK;
    int rc;
    if ((rc = gram_consume(c, score_rows, row_off, pool))) return rc;  // its row buffers are reused
    hipStream_t st = c->fstream;
    for (int h = 0; h < G->ngroups; ++h) HIP_TRY(hipStreamWaitEvent(st, G->stage_ev[h], 0));
    const bool async = !x_out && (int)G->pend.size() <= G->nhrows;
    // without x_out the contraction runs on the R coordinate rows of each locus in two BF16
    // parts (k_score takes each block back to states); x_out needs the q state rows, formed
    // from three parts (FP32 products: the parameters to about 1e-8 relative)
    const int coords = x_out ? 0 : 1;
    const int nmem = (int)c->members.size();
    const int gm = c->opts.gemm_gm_grad > 0 ? c->opts.gemm_gm_grad : 2;
    const int tn = c->opts.gemm_tn_grad > 0 ? c->opts.gemm_tn_grad : 256;
    const size_t n = (size_t)c->L * c->q * c->q + c->q;
    std::vector<float> xdev(x_out ? G->n_item : 0);
    const std::vector<sdca::GramWs::Pending> &P = G->pend;
    const int64_t pbase = (int64_t)G->phalf * G->pcap;
    HIP_TRY(hipEventRecord(G->fev[0], st));
    for (size_t u0 = 0; u0 < P.size(); u0 += G->nmat) {
        const int nb = (int)std::min<size_t>(G->nmat, P.size() - u0);
        // pending entries of the batch by alignment
        std::vector<int> order(nb);
        for (int u = 0; u < nb; ++u) order[u] = (int)u0 + u;
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return P[x].aln < P[y].aln; });
        std::vector<int64_t> moff(nb), soff(nb);
        std::vector<int32_t> mrow(nb), maln(nb), sloc(nb), sld(nb);
        std::vector<sdca::SlotCmd> mc(nb);
        std::vector<sdca::GemmGroup> groups;
        int row = 0;
        for (int u = 0; u < nb;) {
            const int a = P[order[u]].aln;
            const sdca_ctx *A = c->members[a];
            row = (int)round_up(row, sdca::kTile);
            sdca::Parts parts;
            parts.row[0] = row;
            parts.cmd[0] = u;
            int v = u;
            for (; v < nb && P[order[v]].aln == a; ++v) {
                const sdca::GramWs::Pending &pe = P[order[v]];
                moff[v] = (pbase + order[v]) * G->nvs;
                mrow[v] = row;
                maln[v] = a;
                sloc[v] = pe.r;
                sld[v] = (int32_t)A->Kxp;
                soff[v] = (int64_t)v * G->n_item;
                sdca::SlotCmd &k = mc[v];
                k = idle_cmd();
                k.mode = sdca::MODE_INIT;
                k.locus = pe.r;
                k.self_lo = A->scan_split > 0 ? 0 : A->col_off[pe.r];
                k.self_hi = A->scan_split > 0 ? A->col_off[A->scan_split] : A->col_off[pe.r + 1];
                k.slot = v;
                k.row = row;
                k.level = sdca::LV_FINE;
                k.aln = a;
                k.q = A->q;
                k.Kx = (int)A->Kx;
                k.Kxp = (int)A->Kxp;
                k.gt = 0;
                row += coords ? sdca::gram::rows(A->q) : A->q;
            }
            const int rend = (int)round_up(row, sdca::kTile);
            // scores: two BF16 parts (hi + lo) for every row; x_out: three (FP32 products)
            parts.row[1] = coords ? parts.row[0] : rend;
            parts.row[2] = rend;
            parts.cmd[1] = coords ? parts.cmd[0] : v;
            parts.cmd[2] = parts.cmd[3] = v;
            groups.push_back(make_group(c, a, parts, rend, false, 1, 0, gm, tn));
            if (coords) groups.back().q = sdca::gram::rows(A->q);  // rows per item
            row = rend;
            u = v;
        }
        if (row > G->mrows || (int)groups.size() > nmem) return fail(SDCA_E_STATE, "coupling batch exceeds its buffers");
        const int grid = place_groups(groups.data(), (int)groups.size());
        HIP_TRY(hipMemcpyAsync(G->d_moff, moff.data(), nb * 8, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_soff, soff.data(), nb * 8, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_mrow, mrow.data(), nb * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_maln, maln.data(), nb * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_sloc, sloc.data(), nb * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_sld, sld.data(), nb * 4, hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_mcmds, mc.data(), nb * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st));
        HIP_TRY(hipMemcpyAsync(G->d_mgroups, groups.data(), groups.size() * sizeof(sdca::GemmGroup),
                               hipMemcpyHostToDevice, st));
        // the batch's score rows: at its place among the flush's rows, or at the start
        const int64_t hrow0 = async ? (int64_t)u0 : 0;
        float *drows = G->d_rows + (size_t)hrow0 * G->gL;
        sdca::gram::split(st, G->d_moff, G->d_mrow, G->d_maln, nb, G->nchunk, G->d_galn, G->Pend, G->gBp, nullptr,
                          G->Mhi, G->Mlo, coords ? nullptr : G->Ml3, G->mhpart, coords);
        sdca::kern::gemm_materialize(st, SDCA_PREC_SPLIT_BF16, G->d_mgroups, (int)groups.size(), grid, tn, G->M32,
                                     G->Mhi, G->Mlo, G->Ml3, G->gBp, G->d_aln, G->d_mcmds, G->Mout, G->n_item,
                                     G->mhpart);
        sdca::kern::score(st, G->Mout, G->d_soff, G->d_sld, G->d_sloc, G->d_maln, nb, G->gL, G->gL, G->d_aln, drows,
                          coords);
        HIP_TRY(hipGetLastError());
        if (async) {
            for (int v = 0; v < nb; ++v) G->fout.push_back({P[order[v]].item, P[order[v]].aln, hrow0 + v});
            continue;
        }
        HIP_TRY(hipEventRecord(G->fev[1], st));
        if (score_rows)
            HIP_TRY(hipMemcpyAsync(G->h_rows, G->d_rows, (size_t)nb * G->gL * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        c->timing.score += elapsed(G->fev[0], G->fev[1]);
        for (int v = 0; v < nb; ++v) {
            const sdca::GramWs::Pending &pe = P[order[v]];
            const int La = c->members[pe.aln]->L;
            if (score_rows) {
                const int64_t off = row_off ? row_off[pe.item] : (int64_t)pe.item * La;
                std::memcpy(score_rows + off, G->h_rows + (size_t)v * G->gL, (size_t)La * 4);
            }
            if (x_out) {
                HIP_TRY(hipMemcpy(xdev.data(), G->Mout + (size_t)v * G->n_item, G->n_item * 4, hipMemcpyDeviceToHost));
                dev_to_ref(c, pe.r, xdev.data(), x_out + (size_t)pe.item * n);
            }
        }
        HIP_TRY(hipEventRecord(G->fev[0], st));
    }
    HIP_TRY(hipEventRecord(G->pend_free[G->phalf], st));
    if (async) {
        HIP_TRY(hipEventRecord(G->fev[1], st));
        if (score_rows)
            HIP_TRY(hipMemcpyAsync(G->h_rows, G->d_rows, P.size() * (size_t)G->gL * 4, hipMemcpyDeviceToHost, st));
        G->finflight = true;
    }
    G->pend.clear();
    G->phalf ^= 1;
    return SDCA_OK;
}

// ------------------------------------------------ subspace warm start (Gram formulation)
//
// Each locus is first solved with its coefficients restricted to c = Q a, Q an orthonormal
// basis of the dominant kSubRank-dimensional eigen-subspace of K (sdca_subspace.hip): the
// same host Solver on slot vectors of q x kSubRank coefficients, their images under
// M_r = Q^T K_r Q and the fields, stopping at ||g||_2 < subspace_threshold.  The restricted
// optima (a*, h*) are staged per locus; the full solve then starts each locus at
// c = Q a*, K_r c = K Q a* - U V_r a* (k_gwarm) instead of zero.

// Device memory for the start products of one solve call's loci (subspace warm start).
constexpr double kMaxStartBytes = 64.0 * (1ull << 30);

bool subspace_wanted(const sdca_ctx *c)
{
    if (c->opts.subspace_rank <= 0 || c->members.size() != 1 || c->scan_split > 0) return false;
    const sdca_ctx *A = c->members[0];
    return A->scan_split == 0 && A->B >= 2 * sdca::kSubRank;
}

// Orthonormalises the kp rows of Q (row stride ld, n entries each) in place: modified
// Gram-Schmidt, twice.
void mgs_rows(std::vector<double> &Q, int kp, int n, int64_t ld)
{
    for (int pass = 0; pass < 2; ++pass)
        for (int i = 0; i < kp; ++i) {
            double *qi = Q.data() + (size_t)i * ld;
            for (int j = 0; j < i; ++j) {
                const double *qj = Q.data() + (size_t)j * ld;
                double r = 0.0;
                for (int b = 0; b < n; ++b) r += qi[b] * qj[b];
                for (int b = 0; b < n; ++b) qi[b] -= r * qj[b];
            }
            double nn = 0.0;
            for (int b = 0; b < n; ++b) nn += qi[b] * qi[b];
            const double inv = nn > 0.0 ? 1.0 / std::sqrt(nn) : 0.0;
            for (int b = 0; b < n; ++b) qi[b] *= inv;
        }
}

// The basis of A's identical-locus counts: power iteration Q <- orth(K Q) (three products
// on the device from a seeded Gaussian start), then K Q and (Q^T K Q)^-1 (Cholesky on the
// host).  Returns SDCA_OK, or SDCA_E_STATE when Q^T K Q is not positive definite.
int ensure_subspace(sdca_ctx *A, hipStream_t st)
{
    constexpr int kp = sdca::kSubRank;
    if (A->sub_rank == kp) return SDCA_OK;
    if (A->sub_rank < 0) return SDCA_E_STATE;  // these counts had no basis: not tried again
    const int B = A->B, Bp = A->Bp;
    std::vector<double> Qt((size_t)kp * Bp, 0.0), KQt((size_t)kp * Bp, 0.0);
    std::mt19937_64 gen(0x5d5ca11);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int i = 0; i < kp; ++i)
        for (int b = 0; b < B; ++b) Qt[(size_t)i * Bp + b] = nd(gen);
    mgs_rows(Qt, kp, B, Bp);
    double *dQt = nullptr, *dY = nullptr;
    int rc;
    if ((rc = dalloc(&dQt, (size_t)kp * Bp)) || (rc = dalloc(&dY, (size_t)kp * Bp))) {
        dfree(dQt);
        return rc;
    }
    auto product = [&](std::vector<double> &out) -> hipError_t {
        hipError_t e = hipMemcpyAsync(dQt, Qt.data(), Qt.size() * sizeof(double), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return e;
        // Y^T[i][b] = sum_b' Qt[i][b'] K[b][b']
        sdca::gram::gemm_splitk(st, A->d_gram, Bp, dQt, Bp, dY, Bp, Bp, kp, Bp, Bp, 0);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(out.data(), dY, out.size() * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        return e;
    };
    hipError_t e = hipSuccess;
    for (int it = 0; it < 3 && e == hipSuccess; ++it) {
        e = product(KQt);
        if (e == hipSuccess) {
            Qt.swap(KQt);
            mgs_rows(Qt, kp, B, Bp);
        }
    }
    if (e == hipSuccess) e = product(KQt);
    dfree(dQt);
    dfree(dY);
    if (e != hipSuccess) return fail(SDCA_E_HIP, "subspace basis failed: %s", hipGetErrorString(e));
    // M = Q^T K Q (symmetrised), its Cholesky factor and inverse
    std::vector<double> M((size_t)kp * kp), Lc((size_t)kp * kp, 0.0), Mi((size_t)kp * kp, 0.0);
    for (int i = 0; i < kp; ++i)
        for (int j = 0; j <= i; ++j) {
            double v = 0.0, w = 0.0;
            for (int b = 0; b < B; ++b) {
                v += Qt[(size_t)i * Bp + b] * KQt[(size_t)j * Bp + b];
                w += Qt[(size_t)j * Bp + b] * KQt[(size_t)i * Bp + b];
            }
            M[(size_t)i * kp + j] = M[(size_t)j * kp + i] = 0.5 * (v + w);
        }
    for (int j = 0; j < kp; ++j) {
        double d = M[(size_t)j * kp + j];
        for (int k = 0; k < j; ++k) d -= Lc[(size_t)j * kp + k] * Lc[(size_t)j * kp + k];
        if (!(d > 0.0)) {
            A->sub_rank = -1;
            return fail(SDCA_E_STATE, "Q^T K Q is not positive definite");
        }
        Lc[(size_t)j * kp + j] = std::sqrt(d);
        for (int i = j + 1; i < kp; ++i) {
            double v = M[(size_t)i * kp + j];
            for (int k = 0; k < j; ++k) v -= Lc[(size_t)i * kp + k] * Lc[(size_t)j * kp + k];
            Lc[(size_t)i * kp + j] = v / Lc[(size_t)j * kp + j];
        }
    }
    for (int col = 0; col < kp; ++col) {  // M^-1 e_col: forward then back substitution
        std::vector<double> y(kp, 0.0), x(kp, 0.0);
        for (int i = 0; i < kp; ++i) {
            double v = i == col ? 1.0 : 0.0;
            for (int k = 0; k < i; ++k) v -= Lc[(size_t)i * kp + k] * y[k];
            y[i] = v / Lc[(size_t)i * kp + i];
        }
        for (int i = kp - 1; i >= 0; --i) {
            double v = y[i];
            for (int k = i + 1; k < kp; ++k) v -= Lc[(size_t)k * kp + i] * x[k];
            x[i] = v / Lc[(size_t)i * kp + i];
        }
        for (int i = 0; i < kp; ++i) Mi[(size_t)i * kp + col] = x[i];
    }
    for (int i = 0; i < kp; ++i)  // exactly symmetric (the kernels read it both ways)
        for (int j = 0; j < i; ++j) Mi[(size_t)i * kp + j] = Mi[(size_t)j * kp + i] =
            0.5 * (Mi[(size_t)i * kp + j] + Mi[(size_t)j * kp + i]);
    std::vector<double> Q((size_t)Bp * kp), KQ((size_t)Bp * kp);
    for (int i = 0; i < kp; ++i)
        for (int b = 0; b < Bp; ++b) {
            Q[(size_t)b * kp + i] = Qt[(size_t)i * Bp + b];
            KQ[(size_t)b * kp + i] = KQt[(size_t)i * Bp + b];
        }
    dfree(A->d_subQ);
    dfree(A->d_subQt);
    dfree(A->d_subKQ);
    dfree(A->d_subKQt);
    dfree(A->d_subMinv);
    if ((rc = dalloc(&A->d_subQ, Q.size())) || (rc = dalloc(&A->d_subQt, Qt.size())) ||
        (rc = dalloc(&A->d_subKQ, KQ.size())) || (rc = dalloc(&A->d_subKQt, KQt.size())) ||
        (rc = dalloc(&A->d_subMinv, Mi.size())))
        return rc;
    HIP_TRY(hipMemcpy(A->d_subQ, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A->d_subQt, Qt.data(), Qt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A->d_subKQ, KQ.data(), KQ.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A->d_subKQt, KQt.data(), KQt.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(A->d_subMinv, Mi.data(), Mi.size() * sizeof(double), hipMemcpyHostToDevice));
    A->sub_rank = kp;
    return SDCA_OK;
}

// Restricted workspace of the Gram workspace (slot vectors, per-slot constants, packed
// operands) and room for K staged optima.
int ensure_sub_ws(sdca_ctx *c, int32_t K)
{
    constexpr int kp = sdca::kSubRank;
    sdca::GramWs *G = c->gram;
    const sdca_ctx *A = c->members[0];
    int rc;
    if (!G->Vr) {
        G->nvs_r = sdca::gram::vec_len(G->gq, kp);
        G->vstride_r = (sdca::gram::kVecPairs + 2 * G->m) * G->nvs_r;
        G->auxstride = 16 * kp + 64;
        G->nchq_r = (int)((sdca::gram::rows(G->gq) * kp + 255) / 256);
        G->nchd_r = sdca::gram::dots_chunks((int64_t)sdca::gram::rows(G->gq) * kp);
        const int64_t ncols = G->colcap * G->ngroups;
        const int npad = G->gs * G->ngroups;
        if ((rc = dalloc(&G->Vr, (size_t)G->slots * G->vstride_r)) || (rc = dalloc(&G->aux, (size_t)G->slots * G->auxstride)) ||
            (rc = dalloc(&G->Ap, (size_t)ncols * kp)) ||
            (rc = dalloc(&G->Zr, (size_t)ncols * kp * std::max<int64_t>(2, (G->gBp + sdca::kSubKc - 1) / sdca::kSubKc))) ||
            (rc = dalloc(&G->Wc, (size_t)npad * 64)) || (rc = dalloc(&G->d_galn_r, 1)))
            return rc;
    }
    // restricted groups: each of the full solve's groups halved (slots and columns) when
    // four fit, so that two host threads keep four short rounds in flight
    G->rng = G->ngroups;
    if (G->ngroups >= 2 && 2 * G->ngroups <= sdca::kMaxGroups && G->slots >= 256) {
        const int rgs = (G->slots + 2 * G->ngroups - 1) / (2 * G->ngroups);
        if (round_up((int64_t)rgs * G->gq, sdca::gram::kGemmCols) <= G->colcap / 2) G->rng = 2 * G->ngroups;
    }
    G->rgs = (G->slots + G->rng - 1) / G->rng;
    G->rcc = G->colcap * G->ngroups / G->rng;
    // the restricted descriptor: vectors of q x kp coefficients, every "sequence" valid, no
    // own-column term in the products pass (the gradient kernel subtracts it)
    sdca::GramAln g{};
    g.K = nullptr;
    g.statesT = A->d_statesT;
    g.w = A->d_w64;
    g.B = kp;
    g.Bp = kp;
    g.q = A->q;
    g.own = 0;
    g.lamJ2 = 2.0 * A->lambda_J;
    g.lamh2 = 2.0 * A->lambda_h;
    HIP_TRY(hipMemcpy(G->d_galn_r, &g, sizeof(g), hipMemcpyHostToDevice));
    // staged optima: the coordinate rows of every item (the start product reads whole tiles
    // of kGemmCols rows), then 8 fields per item; their start products
    const int R = sdca::gram::rows(A->q);
    const int64_t wrows = round_up((int64_t)std::max(K, 1) * R, sdca::gram::kGemmCols) + sdca::gram::kGemmCols;
    G->wfld = wrows * kp;
    if (G->wcap < K) {
        dfree(G->Wst);
        dfree(G->Cst);
        dfree(G->Kst);
        G->wcap = 0;
        const size_t nst = (size_t)std::max(K, 1) * R * A->Bp;
        if ((rc = dalloc(&G->Wst, (size_t)G->wfld + (size_t)std::max(K, 1) * 8)) || (rc = dalloc(&G->Cst, nst)) ||
            (rc = dalloc(&G->Kst, nst)))
            return rc;
        G->wcap = K;
    }
    return SDCA_OK;
}

// One restricted round of slot group g (the layout of gram_launch_round; the packed
// energies and beliefs reuse the Gram workspace's Zp and Gp).
int sub_launch_round(sdca_ctx *c, int g, const std::vector<sdca::SlotCmd> &cmds)
{
    constexpr int kp = sdca::kSubRank;
    sdca::GramWs *G = c->gram;
    const sdca_ctx *A = c->members[0];
    const int m = G->m, W = sdca::ResLayout{m}.width();
    const int s0 = g * G->rgs, s1 = std::min<int>((int)cmds.size(), s0 + G->rgs);
    std::vector<int> &lst = G->cmd_slot[g];
    lst.clear();
    sdca::SlotCmd *hc = G->h_cmds + s0;
    int col = 0, ninit = 0;
    // INIT commands first: the per-slot constants are formed for them alone
    for (int pass = 0; pass < 2; ++pass)
        for (int i = s0; i < s1; ++i) {
            if (cmds[i].mode == sdca::MODE_IDLE || (cmds[i].mode == sdca::MODE_INIT) != (pass == 0)) continue;
            sdca::SlotCmd &k = hc[lst.size()];
            k = cmds[i];
            k.slot = i;
            k.row = col;
            k.q = A->q;
            lst.push_back(i);
            col += sdca::gram::rows(A->q);
            ninit += pass == 0;
        }
    const int nc = (int)lst.size();
    if (nc == 0) return SDCA_OK;
    hipStream_t st = c->gstream[g];
    hipEvent_t *ev = G->ev[g];
    sdca::SlotCmd *dc = G->d_cmds + s0;
    double *Gp = G->Gp + (size_t)g * G->rcc * G->ldG, *Ep = G->Zp + (size_t)g * G->rcc * G->ldG;
    double *Ap = G->Ap + (size_t)g * G->rcc * kp, *Zr = G->Zr + (size_t)g * G->rcc * kp;
    double *Wc = G->Wc + (size_t)s0 * 64;
    double *upd = G->upd_part + (size_t)s0 * (G->nchq + 1) * 3, *sm = G->sm_part + (size_t)s0 * G->nchunk * 65;
    double *dot = G->dot_part + (size_t)s0 * G->nchd * sdca::ResLayout{m}.ndots(), *res = G->res + (size_t)s0 * W;
    HIP_TRY(hipMemcpyAsync(dc, hc, nc * sizeof(sdca::SlotCmd), hipMemcpyHostToDevice, st));
    HIP_TRY(hipEventRecord(ev[4], st));
    sdca::sub::init(st, dc, ninit, G->d_galn, A->d_subQ, A->d_subMinv, kp, G->aux, G->auxstride);
    sdca::gram::step_vectors(st, m, dc, nc, G->d_galn_r, G->Vr, G->vstride_r, G->nvs_r, upd, G->nchq_r);
    sdca::sub::pack(st, dc, nc, G->d_galn, G->Vr, G->vstride_r, G->nvs_r, kp, G->aux, G->auxstride, Ap, Wc);
    sdca::gram::gemm_splitk(st, A->d_subKQ, kp, Ap, kp, Ep, G->ldG, A->Bp, col, kp, kp, 0);
    sdca::sub::softmax(st, dc, nc, G->d_galn, G->Vr, G->vstride_r, G->nvs_r, kp, Ep, Wc, Gp, G->ldG, sm, G->nchunk);
    // split-K over the sequences: slice z of the product at Zs + z * zs (Zs after the slot
    // groups' first slices), summed by the gradient kernels into Zr
    const int nz = (A->Bp + sdca::kSubKc - 1) / sdca::kSubKc;
    const int64_t zs = G->colcap * G->ngroups * (int64_t)kp;
    sdca::gram::gemm_splitk(st, A->d_subKQt, A->Bp, Gp, G->ldG, Zr, kp, kp, col, A->Bp, sdca::kSubKc, zs);
    sdca::sub::grad(st, dc, nc, G->d_galn, G->Vr, G->vstride_r, G->nvs_r, kp, G->aux, G->auxstride, sm, G->nchunk, Zr,
                    nz, zs, A->d_subMinv, col, 2.0 * A->lambda_J);
    sdca::gram::dots(st, m, dc, nc, G->d_galn_r, G->Vr, G->vstride_r, G->nvs_r, Zr, kp, sm, G->nchunk, dot,
                     G->nchd_r, 1);
    sdca::gram::reduce(st, m, nc, upd, dot, sm, G->nchunk, G->nchq_r, G->nchd_r, res);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(ev[5], st));
    HIP_TRY(hipMemcpyAsync(G->h_res + (size_t)s0 * W, res, (size_t)nc * W * sizeof(double), hipMemcpyDeviceToHost,
                           st));
    return SDCA_OK;
}

int sub_finish_round(sdca_ctx *c, int g, std::vector<double> &res)
{
    sdca::GramWs *G = c->gram;
    const std::vector<int> &lst = G->cmd_slot[g];
    if (lst.empty()) return SDCA_OK;
    HIP_TRY(hipStreamSynchronize(c->gstream[g]));
    const sdca::ResLayout R{G->m};
    const int W = R.width();
    const double *hr = G->h_res + (size_t)g * G->rgs * W;
    for (size_t k = 0; k < lst.size(); ++k) {
        double *o = res.data() + (size_t)lst[k] * W;
        std::memcpy(o, hr + k * W, W * sizeof(double));
        o[R.ginf()] = std::sqrt(std::max(0.0, o[R.ginf()]));
        o[R.gJinf()] = std::sqrt(std::max(0.0, o[R.gJinf()]));
    }
    const sdca_ctx *A = c->members[0];
    std::lock_guard<std::mutex> lk(G->mu);
    c->timing.sub += elapsed(G->ev[g][4], G->ev[g][5]);
    c->timing.sub_evals += 1;
    c->timing.flops_sub += 4.0 * A->B * (double)sdca::kSubRank * sdca::gram::rows(A->q) * (double)lst.size();
    return SDCA_OK;
}

// The restricted solves of items [0, K): their optima into G->Wst (coordinates of item i at
// rows [i R, (i + 1) R), fields at wfld + 8 i), their
// evaluations into nev, and whether each may start the full solve into ok.
int sub_solve(sdca_ctx *c, const int32_t *loci, int32_t K, std::vector<int> &nev, std::vector<char> &ok)
{
    constexpr int kp = sdca::kSubRank;
    sdca::GramWs *G = c->gram;
    const int nslots = G->slots, m = G->m;
    const sdca::ResLayout R{m};
    const SolverLimits lim{c->opts.subspace_threshold, 0.0, c->opts.max_iter};
    Solver S(c, nslots, m, true, &lim);
    const int ng = G->rng, gs = G->rgs;
    const int R1 = sdca::gram::rows(c->members[0]->q);  // coordinate rows of a restricted vector half
    std::vector<double> res((size_t)nslots * R.width(), 0.0);
    std::vector<char> done_flag(nslots, 0);
    nev.assign(K, 0);
    ok.assign(K, 0);
    int next = 0, done = 0;
    auto fill = [&](int i) {
        if (next >= K) {
            S.set_idle(i);
            return;
        }
        const int item = next++;
        S.assign(i, item, 0, loci[item], false);
    };
    auto group_end = [&](int g) { return std::min(nslots, (g + 1) * gs); };
    auto group_active = [&](int g) {
        for (int i = g * gs; i < group_end(g); ++i)
            if (S.slot(i).phase != PH_IDLE) return true;
        return false;
    };
    int rc;
    for (int i = 0; i < nslots; ++i) fill(i);
    HostThreads ht(ng, c->device);
    ht.run([&](int g0, int g1, std::string &) -> int {
        int rc;
        SlotPool pool(SlotPool::default_threads());
        bool inflight[sdca::kMaxGroups] = {};
        for (int g = g0; g < g1; ++g) {
            if (!group_active(g)) continue;
            S.mark_live(g * gs, group_end(g));
            if ((rc = sub_launch_round(c, g, S.cmds()))) return rc;
            inflight[g] = true;
        }
        std::vector<int> finished, keep;
        while (!ht.stopped()) {
            bool any = false;
            for (int g = g0; g < g1; ++g) {
                if (!inflight[g]) continue;
                any = true;
                inflight[g] = false;
                if ((rc = sub_finish_round(c, g, res))) return rc;
                finished.clear();
                const auto th = std::chrono::steady_clock::now();
                pool.run(g * gs, group_end(g), [&](int i) {
                    done_flag[i] = S.slot(i).phase != PH_IDLE && S.process(i, res.data() + (size_t)i * R.width());
                });
                for (int i = g * gs; i < group_end(g); ++i)
                    if (done_flag[i]) finished.push_back(i);
                const double hms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th).count();
                hipStream_t st = c->gstream[g];
                // restricted optima (coordinates, then the fields) into the staged starts
                int64_t *meta = G->h_stage + (size_t)g * gs * 5;
                keep.clear();
                for (int i : finished)
                    if (S.slot(i).status != SDCA_ST_NONFINITE) keep.push_back(i);
                const int nk = (int)keep.size();
                for (int k = 0; k < nk; ++k) {
                    const Slot &sl = S.slot(keep[k]);
                    meta[k] = (int64_t)keep[k] * G->vstride_r +
                              (sl.from_xt ? sdca::gram::kVecXt : sdca::gram::kVecXc) * G->nvs_r;
                    meta[nk + k] = (int64_t)sl.item * R1 * kp;
                    meta[2 * nk + k] = (int64_t)R1 * kp;
                    meta[3 * nk + k] = 2 * (int64_t)R1 * kp;
                    meta[4 * nk + k] = G->wfld + (int64_t)sl.item * 8 - (int64_t)sl.item * R1 * kp;
                    ok[sl.item] = 1;
                }
                if (nk > 0) {
                    int64_t *dmeta = G->d_stage + (size_t)g * gs * 5;
                    HIP_TRY(hipMemcpyAsync(dmeta, meta, (size_t)nk * 5 * sizeof(int64_t), hipMemcpyHostToDevice, st));
                    sdca::gram::gather(st, G->Vr, G->Wst, dmeta, nk, (int64_t)R1 * kp, 8);
                    HIP_TRY(hipGetLastError());
                }
                {
                    std::lock_guard<std::mutex> lk(G->mu);
                    c->timing.host += hms;
                    for (int i : finished) {
                        nev[S.slot(i).item] = S.slot(i).nfeval;
                        ++done;
                        fill(i);
                    }
                }
                if (group_active(g)) {
                    S.mark_live(g * gs, group_end(g));
                    if ((rc = sub_launch_round(c, g, S.cmds()))) return rc;
                    inflight[g] = true;
                }
            }
            if (!any) break;
        }
        return SDCA_OK;
    });
    if ((rc = ht.result())) return fail(rc, "%s", ht.message().c_str());
    for (int g = 0; g < ng; ++g) HIP_TRY(hipStreamSynchronize(c->gstream[g]));
    return SDCA_OK;
}

int gram_solve_batch(sdca_ctx *c, const int32_t *alns, const int32_t *loci, int32_t K, float *score_rows,
                     const int64_t *row_off, double *x_out, sdca_locus_stats *st_out)
{
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    for (sdca_ctx *A : c->members)
        if ((rc = ensure_gram_matrix(A, c->stream))) return rc;
    const int nslots = slots_for(c->opts, true, K);
    const int m = history_for(c->opts, true);
    if ((rc = ensure_gram_ws(c, nslots, m))) return rc;
    if ((rc = upload_gram_alns(c))) return rc;
    const sdca::ResLayout R{m};
    sdca::GramWs *G = c->gram;
    // nothing of an earlier call that returned early: its flush in flight, its pending loci
    if (G->finflight) {
        HIP_TRY(hipStreamSynchronize(c->fstream));
        G->finflight = false;
        G->fout.clear();
    }
    G->pend.clear();
    // subspace warm start: the restricted solves of every listed locus first
    std::vector<int> sub_nev;
    std::vector<char> sub_ok;
    G->warm_rounds = false;
    // (the start products take 16 R Bp bytes per locus: batches beyond kMaxStartBytes start cold)
    const bool sub_fits = (double)K * 16.0 * sdca::gram::rows(c->members[0]->q) * c->members[0]->Bp <= kMaxStartBytes;
    if (K > 0 && sub_fits && subspace_wanted(c) && ensure_subspace(c->members[0], c->stream) == SDCA_OK) {
        if ((rc = ensure_sub_ws(c, K))) return rc;
        if ((rc = sub_solve(c, loci, K, sub_nev, sub_ok))) return rc;
        // the start products of every restricted optimum at once (loci whose restricted solve
        // failed start from zero: their columns are not read)
        const sdca_ctx *A = c->members[0];
        sdca::sub::starts(c->gstream[0], A->d_subQ, A->d_subKQ, sdca::kSubRank, A->Bp, G->Wst,
                          (int64_t)K * sdca::gram::rows(A->q), G->Cst, G->Kst);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(c->gstream[0]));
        G->warm_rounds = true;
    }
    G->sub_nev = sub_nev.empty() ? nullptr : sub_nev.data();
    Solver S(c, nslots, m, true);
    const int ng = G->ngroups, gs = G->gs;
    std::vector<double> res((size_t)nslots * R.width(), 0.0);
    std::vector<char> done_flag(nslots, 0);
    // Order of the full solves.  After the restricted solves, loci start in decreasing order
    // of their restricted evaluations, which predict the full ones (C4: correlation 0.8):
    // the longest solves start first, so the last rounds are not left to a few slow loci
    // (longest-processing-time-first scheduling).  Every locus' result is the same in any
    // order: its rounds depend on its own slot only.
    std::vector<int> order(K);
    std::iota(order.begin(), order.end(), 0);
    if (G->warm_rounds)
        std::stable_sort(order.begin(), order.end(), [&](int u, int v) { return sub_nev[u] > sub_nev[v]; });
    int next = 0, done = 0, noconv = 0, nonfinite = 0;
    auto fill = [&](int i) {
        if (next >= K) {
            S.set_idle(i);
            return;
        }
        const int item = order[next++];
        S.assign(i, item, alns ? alns[item] : 0, loci[item], G->warm_rounds && sub_ok[item]);
    };
    auto group_end = [&](int g) { return std::min(nslots, (g + 1) * gs); };
    auto group_active = [&](int g) {
        for (int i = g * gs; i < group_end(g); ++i)
            if (S.slot(i).phase != PH_IDLE) return true;
        return false;
    };
    for (int i = 0; i < nslots; ++i) fill(i);
    int64_t rounds = 0;
    // The groups' rounds alternate: while the host steps one group's line searches (and
    // forms its finished loci's couplings), the other group's round runs on the device.
    // With four groups, two host threads step two groups each.
    HostThreads ht(ng, c->device);
    ht.run([&](int g0, int g1, std::string &msg) -> int {
        int rc;
        SlotPool pool(SlotPool::default_threads());
        bool inflight[sdca::kMaxGroups] = {};
        for (int g = g0; g < g1; ++g) {
            if (!group_active(g)) continue;
            S.mark_live(g * gs, group_end(g));
            if ((rc = gram_launch_round(c, g, S.cmds()))) return (msg = g_err, rc);
            inflight[g] = true;
        }
        std::vector<int> finished;
        while (!ht.stopped()) {
            bool any = false;
            for (int g = g0; g < g1; ++g) {
                if (!inflight[g]) continue;
                any = true;
                inflight[g] = false;
                if ((rc = gram_finish_round(c, g, S.cmds(), res))) return (msg = g_err, rc);
                finished.clear();
                const auto th = std::chrono::steady_clock::now();
                pool.run(g * gs, group_end(g), [&](int i) {
                    done_flag[i] = S.slot(i).phase != PH_IDLE && S.process(i, res.data() + (size_t)i * R.width());
                });
                for (int i = g * gs; i < group_end(g); ++i)
                    if (done_flag[i]) finished.push_back(i);
                const double hms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th).count();
                {
                    std::lock_guard<std::mutex> lk(G->mu);
                    c->timing.host += hms;
                    ++rounds;
                    if (!finished.empty()) {
                        if ((rc = gram_stage(c, g, S, finished, st_out))) return (msg = g_err, rc);
                        for (int i : finished) {
                            const Slot &sl = S.slot(i);
                            noconv += (sl.status == SDCA_ST_ITER_CAP || sl.status == SDCA_ST_LS_FAIL);
                            nonfinite += (sl.status == SDCA_ST_NONFINITE);
                            ++done;
                            fill(i);
                        }
                        if (c->opts.verbose)
                            fprintf(stderr, "sdca: %d / %d loci solved after %lld rounds\n", done, K, (long long)rounds);
                    }
                    // a full batch of finished loci: couplings and score rows (on their own
                    // stream, while the rounds go on)
                    if ((rc = gram_flush(c, false, score_rows, row_off, x_out, &pool))) return (msg = g_err, rc);
                }
                if (group_active(g)) {
                    S.mark_live(g * gs, group_end(g));
                    if ((rc = gram_launch_round(c, g, S.cmds()))) return (msg = g_err, rc);
                    inflight[g] = true;
                }
            }
            if (!any) break;
        }
        return SDCA_OK;
    });
    if ((rc = ht.result())) return fail(rc, "%s", ht.message().c_str());
    {
        SlotPool pool(SlotPool::default_threads());
        if ((rc = gram_flush(c, true, score_rows, row_off, x_out, &pool))) return rc;
        if ((rc = gram_consume(c, score_rows, row_off, &pool))) return rc;
    }
    G->warm_rounds = false;
    G->sub_nev = nullptr;
    c->timing.total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (nonfinite) return fail(SDCA_E_NONFINITE, "%d loci reached a non-finite objective", nonfinite);
    if (noconv) return fail(SDCA_E_NOCONV, "%d loci stopped without reaching the gradient threshold", noconv);
    return SDCA_OK;
}

// One-hot int8 X8 (L*QP rows, b contiguous) and the marginal counts, built on first use.
int ensure_x8(sdca_ctx *c)
{
    if (c->d_X8) return SDCA_OK;
    c->QP = c->q <= 4 ? 4 : 8;
    const int64_t Kc = (int64_t)c->L * c->QP;
    c->Kc_pad = round_up(Kc, sdca::kTile);
    c->Kc_alloc = round_up(Kc, 1024) + 1024;  // a zero panel past the end for unaligned panels
    int rc;
    if ((rc = dalloc(&c->d_X8, (size_t)c->Kc_alloc * c->Bp))) return rc;
    if ((rc = dalloc(&c->d_cnt, (size_t)c->Kc_alloc))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    sdca::kern::build_x8(c->stream, c->d_states, c->B, c->L, c->QP, c->Bp, c->d_X8);
    sdca::kern::marginals(c->stream, c->d_states, c->B, c->L, c->QP, c->d_cnt);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(c->stream));
    return SDCA_OK;
}

// The streams and events of a context on its device (default priority; blocking streams,
// so they are ordered after the null-stream zero fills of dalloc()).  One stream per slot
// group; group 0 uses c->stream.
int create_streams(sdca_ctx *c)
{
    if (hipSetDevice(c->device) != hipSuccess) return fail(SDCA_E_HIP, "hipSetDevice(%d) failed", c->device);
    if (hipStreamCreate(&c->stream) != hipSuccess) return fail(SDCA_E_HIP, "hipStreamCreate failed");
    for (auto &e : c->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(SDCA_E_HIP, "hipEventCreate failed");
    c->gstream[0] = c->stream;
    for (int g = 1; g < sdca::kMaxGroups; ++g)
        if (hipStreamCreate(&c->gstream[g]) != hipSuccess) return fail(SDCA_E_HIP, "hipStreamCreate failed");
    if (hipStreamCreate(&c->fstream) != hipSuccess) return fail(SDCA_E_HIP, "hipStreamCreate failed");
    for (auto &g : c->gev)
        for (auto &e : g)
            if (hipEventCreate(&e) != hipSuccess) return fail(SDCA_E_HIP, "hipEventCreate failed");
    return SDCA_OK;
}

int check_ctx(sdca_ctx *c)
{
    if (!c) return fail(SDCA_E_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    return SDCA_OK;
}

}  // namespace

// ====================================================================== C ABI

extern "C" {

const char *sdca_last_error(void) { return g_err.c_str(); }

void sdca_default_opts(sdca_opts *o)
{
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->precision = SDCA_PREC_FP32;
    o->batch_loci = 128;
    o->history = 0;
    o->max_iter = 2000;
    o->max_ls = 20;
    o->grad_threshold = 1e-3;
    o->verbose = 0;
    o->autotune = 1;
    o->coarse_factor = 4;
    o->fine_factor = 10;
    o->j_threshold = 0.0;
    o->subspace_rank = 0;
    o->subspace_threshold = 3e-2;
}

int sdca_ctx_create(sdca_ctx **out, const uint8_t *states, int32_t B, int32_t L, int32_t q, int32_t gap,
                    const sdca_opts *opts)
{
    if (!out) return fail(SDCA_E_ARG, "out is null");
    *out = nullptr;
    if (!states || B < 1 || L < 1) return fail(SDCA_E_ARG, "need states, B >= 1 and L >= 1 (got B=%d L=%d)", B, L);
    if (q < 2 || q > 8) return fail(SDCA_E_ARG, "q_env must be in [2, 8] (got %d)", q);
    if (gap >= q) return fail(SDCA_E_ARG, "gap_state %d outside [0, %d)", gap, q);
    sdca_opts o;
    if (opts) o = *opts;
    else sdca_default_opts(&o);
    if (o.gemm_ksplit < 0 || o.gemm_ksplit > 4096 || o.gemm_gm_energy < 0 || o.gemm_gm_energy > 64 ||
        o.gemm_gm_grad < 0 || o.gemm_gm_grad > 64)
        return fail(SDCA_E_ARG, "gemm_ksplit must be in [0, 4096] and gemm_gm_* in [0, 64]");
    for (int tn : {o.gemm_tn_energy, o.gemm_tn_grad})
        if (tn != 0 && tn != 128 && tn != 256) return fail(SDCA_E_ARG, "gemm_tn_* must be 0, 128 or 256");
    if (o.coarse_factor < 0 || o.coarse_factor > 1000000)
        return fail(SDCA_E_ARG, "coarse_factor must be 0 (off) or a factor >= 1 (got %d)", o.coarse_factor);
    if (o.fine_factor < 0 || o.fine_factor > 1000000)
        return fail(SDCA_E_ARG, "fine_factor must be 0 (no fine level) or a factor >= 1 (got %d)", o.fine_factor);
    if (!(o.j_threshold >= 0.0)) return fail(SDCA_E_ARG, "j_threshold must be >= 0");
    if (o.pipeline_groups < 0 || o.pipeline_groups > sdca::kMaxGroups || o.pipeline_groups == 3)
        return fail(SDCA_E_ARG, "pipeline_groups must be 0, 1, 2 or 4");
    if (o.precision != SDCA_PREC_FP32 && o.precision != SDCA_PREC_SPLIT_BF16)
        return fail(SDCA_E_ARG, "unknown precision %d", o.precision);
    if (o.formulation < SDCA_FORM_AUTO || o.formulation > SDCA_FORM_GRAM)
        return fail(SDCA_E_ARG, "unknown formulation %d", o.formulation);
    if (!valid_history(o.history)) return fail(SDCA_E_ARG, "history must be 0 (the formulation's default) or one of 1, 3, 5, 8, 10, 16, 24, 32 (got %d)", o.history);
    if (o.batch_loci < 0) return fail(SDCA_E_ARG, "batch_loci must be >= 1 (or 0: the formulation's default)");
    if (o.max_iter < 1 || o.max_ls < 1) return fail(SDCA_E_ARG, "max_iter and max_ls must be >= 1");
    if (!(o.grad_threshold >= 0.0)) return fail(SDCA_E_ARG, "grad_threshold must be >= 0");
    if (o.subspace_rank != 0 && o.subspace_rank != sdca::kSubRank)
        return fail(SDCA_E_ARG, "subspace_rank must be 0 (off) or %d (got %d)", sdca::kSubRank, o.subspace_rank);
    if (o.subspace_rank && !(o.subspace_threshold > 0.0)) return fail(SDCA_E_ARG, "subspace_threshold must be > 0");
    for (int64_t e = 0; e < (int64_t)B * L; ++e)
        if (states[e] >= q) return fail(SDCA_E_ARG, "state %d at element %lld outside [0, %d)", states[e], (long long)e, q);

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SDCA_E_HIP, "no HIP device available");
    if (o.device < 0 || o.device >= ndev) return fail(SDCA_E_ARG, "device %d not present (%d devices)", o.device, ndev);

    auto *c = new sdca_ctx();
    c->opts = o;
    c->device = o.device;
    c->B = B;
    c->L = L;
    c->q = q;
    c->gap = gap;
    c->h_states.assign(states, states + (size_t)B * L);
    // observed states per locus -> columns of X
    c->col_off.assign(L + 1, 0);
    c->state_col.assign((size_t)L * q, -1);
    std::vector<uint8_t> seen((size_t)L * q, 0);
    for (int b = 0; b < B; ++b)
        for (int j = 0; j < L; ++j) seen[(size_t)j * q + states[(size_t)b * L + j]] = 1;
    int64_t col = 0;
    for (int j = 0; j < L; ++j) {
        c->col_off[j] = (int32_t)col;
        for (int a = 0; a < q; ++a)
            if (seen[(size_t)j * q + a]) c->state_col[(size_t)j * q + a] = (int32_t)col++;
    }
    c->col_off[L] = (int32_t)col;
    c->Kx = col;
    c->Kxp = round_up(c->Kx + 1, sdca::kTileWide);
    c->Bp = (int)round_up(B, sdca::kTileWide);
    if (c->Kxp >= (int64_t(1) << 31) / 8) {
        delete c;
        return fail(SDCA_E_ARG, "alignment too wide (K_x = %lld)", (long long)col);
    }

    int rc = SDCA_OK;
    auto bail = [&](int code) {
        delete c;
        return code;
    };
    if ((rc = create_streams(c))) return bail(rc);
    c->members.assign(1, c);
    std::vector<uint8_t> sT((size_t)B * L);
    for (int b = 0; b < B; ++b)
        for (int j = 0; j < L; ++j) sT[(size_t)j * B + b] = states[(size_t)b * L + j];
    if ((rc = dalloc(&c->d_states, (size_t)B * L))) return bail(rc);
    if ((rc = dalloc(&c->d_statesT, (size_t)B * L))) return bail(rc);
    if ((rc = dalloc(&c->d_state_col, (size_t)L * q))) return bail(rc);
    if ((rc = dalloc(&c->d_w, (size_t)c->Bp))) return bail(rc);
    if ((rc = dalloc(&c->d_X, (size_t)c->Bp * (c->Kxp / 32)))) return bail(rc);
    if ((rc = dalloc(&c->d_Xt, (size_t)c->Kxp * (c->Bp / 32)))) return bail(rc);
    if (hipMemcpy(c->d_states, states, (size_t)B * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_statesT, sT.data(), (size_t)B * L, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_state_col, c->state_col.data(), c->state_col.size() * 4, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(SDCA_E_HIP, "upload of the alignment failed"));
    if (hipDeviceSynchronize() != hipSuccess) return bail(fail(SDCA_E_HIP, "zero fill of X failed"));
    sdca::kern::build_x(c->stream, c->d_states, B, L, q, c->d_state_col, (int)c->Kx, c->Bp, (int)c->Kxp, c->d_X,
                        c->d_Xt);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
        return bail(fail(SDCA_E_HIP, "building X failed"));
    *out = c;
    return SDCA_OK;
}

void sdca_ctx_destroy(sdca_ctx *ctx) { delete ctx; }

int64_t sdca_num_columns(const sdca_ctx *ctx) { return ctx ? ctx->Kx : -1; }

int sdca_set_scan_split(sdca_ctx *c, int32_t n_first)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (n_first < 0 || n_first >= c->L)
        return fail(SDCA_E_ARG, "the first alignment's loci must number in [0, L) (got %d of L=%d)", n_first, c->L);
    c->scan_split = n_first;
    return SDCA_OK;
}

int sdca_reweight(sdca_ctx *c, double theta, const uint32_t *mult, double *w_out, uint64_t *int_w_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!w_out) return fail(SDCA_E_ARG, "w_out is null");
    const int B = c->B, L = c->L, q = c->q;
    const int W = (L + 63) / 64;
    const int64_t dmax = sdca::identity_dmax(L, theta);
    uint64_t *d_planes = nullptr;
    unsigned long long *d_iw = nullptr;
    uint32_t *d_mult = nullptr;
    std::vector<unsigned long long> iw(B);
    for (int i = 0; i < B; ++i) iw[i] = mult ? mult[i] : 1u;
    auto cleanup = [&]() {
        dfree(d_planes);
        dfree(d_iw);
        dfree(d_mult);
    };
    if ((rc = dalloc(&d_planes, (size_t)B * q * W)) || (rc = dalloc(&d_iw, (size_t)B))) {
        cleanup();
        return rc;
    }
    if (mult) {
        if ((rc = dalloc(&d_mult, (size_t)B))) {
            cleanup();
            return rc;
        }
        (void)hipMemcpy(d_mult, mult, (size_t)B * 4, hipMemcpyHostToDevice);
    }
    (void)hipMemcpy(d_iw, iw.data(), (size_t)B * 8, hipMemcpyHostToDevice);
    hipEvent_t e0 = c->ev[6], e1 = c->ev[7];
    (void)hipEventRecord(e0, c->stream);
    sdca::kern::bitplanes(c->stream, c->d_states, B, L, q, W, d_planes);
    sdca::kern::hamming_neighbours(c->stream, d_planes, B, q, W, L, dmax, d_mult, d_iw);
    (void)hipEventRecord(e1, c->stream);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(iw.data(), d_iw, (size_t)B * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    cleanup();
    if (e != hipSuccess) return fail(SDCA_E_HIP, "reweighting kernels failed: %s", hipGetErrorString(e));
    c->timing.reweight += elapsed(e0, e1);
    for (int i = 0; i < B; ++i) {
        const double mi = mult ? (double)mult[i] : 1.0;
        w_out[i] = mi / (double)iw[i];  // :95
        if (int_w_out) int_w_out[i] = iw[i];
    }
    return SDCA_OK;
}

int sdca_set_weights(sdca_ctx *c, const double *w, double lambda_h_user, double lambda_J_user, double *out3)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!w) return fail(SDCA_E_ARG, "weights are null");
    double beff = 0.0;
    for (int i = 0; i < c->B; ++i) {
        if (!std::isfinite(w[i]) || w[i] < 0.0) return fail(SDCA_E_NONFINITE, "weight %d is %g", i, w[i]);
        beff += w[i];  // std::accumulate from 0.0, plmDCA_optimizer_parameters.hpp:261
    }
    const double auto_lambda = (beff > 500.0) ? 0.1 : 1.0 - ((1.0 - 0.1) * beff / 500.0);  // :263
    c->lambda_J = (lambda_J_user < 0.0) ? auto_lambda * beff / 2.0 : lambda_J_user * beff;  // :265-274
    c->lambda_h = (lambda_h_user < 0.0) ? auto_lambda * beff : lambda_h_user * beff;        // :276-284
    c->B_eff = beff;
    c->h_w.assign(w, w + c->B);
    std::vector<float> wf(c->Bp, 0.f);
    for (int i = 0; i < c->B; ++i) wf[i] = (float)w[i];
    HIP_TRY(hipMemcpy(c->d_w, wf.data(), (size_t)c->Bp * 4, hipMemcpyHostToDevice));
    c->have_weights = true;
    if (out3) {
        out3[0] = beff;
        out3[1] = c->lambda_h;
        out3[2] = c->lambda_J;
    }
    return SDCA_OK;
}

int sdca_eval_batch(sdca_ctx *c, const int32_t *loci, int32_t K, const double *x, double *f_out, double *g_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (!c->have_weights) return fail(SDCA_E_STATE, "sdca_set_weights must be called first");
    if (c->scan_split > 0) return fail(SDCA_E_STATE, "sdca_eval_batch does not evaluate the scan objective");
    if (K < 0 || (K > 0 && (!loci || !x || !f_out || !g_out))) return fail(SDCA_E_ARG, "bad arguments");
    for (int k = 0; k < K; ++k)
        if (loci[k] < 0 || loci[k] >= c->L) return fail(SDCA_E_ARG, "locus %d out of range", loci[k]);
    if (K == 0) return SDCA_OK;
    const int nslots = slots_for(c->opts, false, K);
    if (history_for(c->opts, false) > 16)
        return fail(SDCA_E_ARG, "history %d: histories above 16 are supported by the Gram formulation only",
                    c->opts.history);
    if ((rc = ensure_workspace(c, nslots, history_for(c->opts, false)))) return rc;
    if ((rc = upload_alns(c))) return rc;
    if (!c->d_X0) {
        if ((rc = dalloc(&c->d_X0, (size_t)nslots * c->n_slot))) return rc;
        HIP_TRY(hipDeviceSynchronize());
    }
    const int m = c->ws_m;
    const size_t n = (size_t)c->L * c->q * c->q + c->q;
    std::vector<float> stage((size_t)nslots * c->n_slot), gdev(c->n_slot);
    std::vector<double> unobs(nslots);
    std::vector<sdca::SlotCmd> cmds(nslots, idle_cmd());
    for (int k0 = 0; k0 < K; k0 += nslots) {
        const int cnt = std::min(nslots, K - k0);
        for (int i = 0; i < nslots; ++i) {
            cmds[i] = idle_cmd();
            if (i >= cnt) continue;
            const int r = loci[k0 + i];
            unobs[i] = ref_to_dev(c, r, x + (size_t)(k0 + i) * n, stage.data() + (size_t)i * c->n_slot);
            cmds[i].mode = sdca::MODE_INIT;
            cmds[i].locus = r;
            cmds[i].gt = 0;
            cmds[i].gk = -1;
            cmds[i].pnew = -1;
            cmds[i].self_lo = c->col_off[r];
            cmds[i].self_hi = c->col_off[r + 1];
            cmds[i].init_src = 1.f;
            cmds[i].level = c->opts.j_threshold > 0.0 && (c->opts.precision == SDCA_PREC_FP32 || c->opts.fine_factor > 0)
                                ? sdca::LV_FINE
                                : sdca::LV_SPLIT;
        }
        HIP_TRY(hipMemcpy(c->d_X0, stage.data(), (size_t)cnt * c->n_slot * 4, hipMemcpyHostToDevice));
        for (int g = 0; g < c->ngroups; ++g)
            if ((rc = launch_round(c, g, cmds.data() + (size_t)g * c->gslots))) return rc;
        for (int g = 0; g < c->ngroups; ++g)
            if ((rc = finish_round(c, g, cmds.data() + (size_t)g * c->gslots))) return rc;
        const sdca::ResLayout R{m};
        for (int i = 0; i < cnt; ++i) {
            const int r = loci[k0 + i];
            const double *rr = c->h_res + (size_t)i * R.width();
            f_out[k0 + i] = rr[R.fdata()] + c->lambda_J * rr[R.xsqJ()] + c->lambda_h * rr[R.xsqh()] +
                            c->lambda_J * unobs[i];
            HIP_TRY(hipMemcpy(gdev.data(), c->d_gbuf + (size_t)i * 2 * c->n_slot, c->n_slot * 4,
                              hipMemcpyDeviceToHost));  // gradient buffer 0
            double *g = g_out + (size_t)(k0 + i) * n;
            dev_to_ref(c, r, gdev.data(), g);
            // states never observed at j: gradient 2 lambda_J x (plmDCA_objective.hpp:217)
            const double *xr = x + (size_t)(k0 + i) * n;
            for (int j = 0; j < c->L; ++j) {
                if (j == r) continue;
                for (int a = 0; a < c->q; ++a) {
                    if (c->state_col[(size_t)j * c->q + a] >= 0) continue;
                    const size_t o = sdca::ref_index(c->L, c->q, j, a, 0);
                    for (int s = 0; s < c->q; ++s) g[o + s] = 2.0 * c->lambda_J * xr[o + s];
                }
            }
        }
    }
    return SDCA_OK;
}

int sdca_solve_loci(sdca_ctx *c, const int32_t *loci, int32_t K, const double *warm, float *score_rows,
                    double *x_out, sdca_locus_stats *st)
{
    return sdca_solve_impl(c, loci, K, warm, nullptr, score_rows, x_out, st);
}

}  // extern "C"

int sdca_solve_impl(sdca_ctx *c, const int32_t *loci, int32_t K, const double *warm, const sdca::CompactWarm *cw,
                    float *score_rows, double *x_out, sdca_locus_stats *st_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    return sdca_solve_batch(c, nullptr, loci, K, warm, cw, score_rows, nullptr, x_out, st_out);
}

int sdca_solve_batch(sdca_ctx *c, const int32_t *alns, const int32_t *loci, int32_t K, const double *warm,
                     const sdca::CompactWarm *cw, float *score_rows, const int64_t *row_off, double *x_out,
                     sdca_locus_stats *st_out)
{
    int rc;
    const int nmem = (int)c->members.size();
    for (const sdca_ctx *A : c->members)
        if (!A->have_weights) return fail(SDCA_E_STATE, "sdca_set_weights must be called first");
    if (K < 0 || (K > 0 && !loci)) return fail(SDCA_E_ARG, "bad arguments");
    if (nmem != 1 && (warm || cw || x_out)) return fail(SDCA_E_ARG, "warm starts and x_out need a single alignment");
    for (int k = 0; k < K; ++k) {
        const int a = alns ? alns[k] : 0;
        if (a < 0 || a >= nmem) return fail(SDCA_E_ARG, "alignment %d of item %d out of range", a, k);
        const sdca_ctx *A = c->members[a];
        if (loci[k] < 0 || loci[k] >= A->L) return fail(SDCA_E_ARG, "locus %d out of range", loci[k]);
        if (A->scan_split > 0) {
            // the masked first alignment must stay at zero: no start points there
            if (warm || cw) return fail(SDCA_E_STATE, "a two-alignment scan solves from a cold start only");
            if (loci[k] >= A->scan_split)
                return fail(SDCA_E_ARG, "locus %d is not in the first alignment (loci [0, %d))", loci[k],
                            A->scan_split);
        }
    }
    if (K == 0) return SDCA_OK;
    if (c->opts.formulation == SDCA_FORM_GRAM && (warm || cw))
        return fail(SDCA_E_ARG, "the Gram formulation solves from a cold start only");
    if (gram_wanted(c, warm || cw)) return gram_solve_batch(c, alns, loci, K, score_rows, row_off, x_out, st_out);
    if (history_for(c->opts, false) > 16)
        return fail(SDCA_E_ARG, "history %d: histories above 16 are supported by the Gram formulation only",
                    c->opts.history);
    const auto t0 = std::chrono::steady_clock::now();
    const int nslots = slots_for(c->opts, false, K);
    if ((rc = ensure_workspace(c, nslots, history_for(c->opts, false)))) return rc;
    if ((rc = upload_alns(c))) return rc;
    const bool any_warm = warm || cw;
    if (any_warm && !c->d_X0) {
        if ((rc = dalloc(&c->d_X0, (size_t)nslots * c->n_slot))) return rc;
        HIP_TRY(hipDeviceSynchronize());
    }
    int32_t *d_wl = nullptr;  // compact warm starts: slot, item, locus of each refill (per group)
    if (cw) {
        if ((rc = dalloc(&d_wl, (size_t)3 * nslots))) return rc;
        HIP_TRY(hipDeviceSynchronize());
    }
    struct Guard {
        int32_t *&p;
        ~Guard() { dfree(p); }
    } guard{d_wl};
    const int m = c->ws_m;
    const size_t n = (size_t)c->L * c->q * c->q + c->q;
    const sdca::ResLayout R{m};
    const int ng = c->ngroups, gs = c->gslots;
    Solver S(c, nslots, c->ws_m, false);
    std::vector<float> stage(warm ? c->n_slot : 0), xdev(x_out ? c->n_slot : 0);
    std::vector<int32_t> wl_slot[sdca::kMaxGroups], wl_item[sdca::kMaxGroups], wl_locus[sdca::kMaxGroups];
    int next = 0, done = 0, noconv = 0, nonfinite = 0;
    auto fill = [&](int i) -> int {
        if (next >= K) {
            S.set_idle(i);
            return SDCA_OK;
        }
        const int item = next++;
        const int r = loci[item];
        const int a = alns ? alns[item] : 0;
        if (warm) {
            (void)ref_to_dev(c, r, warm + (size_t)item * n, stage.data());
            HIP_TRY(hipMemcpy(c->d_X0 + (size_t)i * c->n_slot, stage.data(), c->n_slot * 4, hipMemcpyHostToDevice));
        }
        if (cw) {
            const int g = i / gs;
            wl_slot[g].push_back(i);
            wl_item[g].push_back(item);
            wl_locus[g].push_back(r);
        }
        S.assign(i, item, a, r, any_warm, cw && cw->as_direction);
        return SDCA_OK;
    };
    // stage the compact warm starts of group g's refills into their slots' X0 vectors
    auto stage_compact = [&](int g) -> int {
        if (!cw || wl_slot[g].empty()) return SDCA_OK;
        hipStream_t st = c->gstream[g];
        const int nw = (int)wl_slot[g].size();
        // the group's pinned list is rewritten only after the group's next finish_round,
        // which waits for this copy
        int32_t *packed = c->h_wl + (size_t)3 * g * gs;
        std::copy(wl_slot[g].begin(), wl_slot[g].end(), packed);
        std::copy(wl_item[g].begin(), wl_item[g].end(), packed + nw);
        std::copy(wl_locus[g].begin(), wl_locus[g].end(), packed + 2 * nw);
        int32_t *dw = d_wl + (size_t)3 * g * gs;
        HIP_TRY(hipMemcpyAsync(dw, packed, (size_t)3 * nw * 4, hipMemcpyHostToDevice, st));
        // blocks: into the staged start, or into the direction buffer (first search
        // direction, in the pair precision)
        const bool dbf = cw->as_direction && c->opts.precision == SDCA_PREC_SPLIT_BF16;
        void *XJ = cw->as_direction ? (void *)c->d_dbuf : (void *)c->d_X0;
        const size_t eb = dbf ? 2 : 4;  // bytes per element of XJ
        const int64_t jstride = c->n_slot;
        for (int i : wl_slot[g]) {
            HIP_TRY(hipMemsetAsync(c->d_X0 + (size_t)i * c->n_slot, 0, c->n_slot * 4, st));
            if (cw->as_direction)
                HIP_TRY(hipMemsetAsync((uint8_t *)XJ + (size_t)i * jstride * eb, 0, c->n_slot * eb, st));
        }
        sdca::kern::scatter_warm(st, dw, dw + nw, dw + 2 * nw, cw->d_partners, cw->d_J, cw->d_h, nw, cw->P, c->q,
                                 c->Kxp, (int)c->Kx, c->d_state_col, c->n_slot, c->d_X0, XJ, dbf, jstride);
        HIP_TRY(hipGetLastError());
        wl_slot[g].clear();
        wl_item[g].clear();
        wl_locus[g].clear();
        return SDCA_OK;
    };
    auto group_active = [&](int g) {
        for (int i = g * gs; i < (g + 1) * gs; ++i)
            if (S.slot(i).phase != PH_IDLE) return true;
        return false;
    };
    for (int i = 0; i < nslots; ++i)
        if ((rc = fill(i))) return rc;
    bool inflight[sdca::kMaxGroups] = {};
    for (int g = 0; g < ng; ++g) {
        if ((rc = stage_compact(g))) return rc;
        if (group_active(g)) {
            S.mark_live(g * gs, (g + 1) * gs);
            if ((rc = launch_round(c, g, S.cmds().data() + (size_t)g * gs))) return rc;
            inflight[g] = true;
        }
    }
    std::vector<int> finished;
    // score rows on their way to the host, per group: the D2H copy into the group's pinned
    // slice runs on the group's stream ahead of its next round; the rows go on to the
    // caller's buffer once that round has been launched
    std::vector<int32_t> row_items[sdca::kMaxGroups];
    bool scoring[sdca::kMaxGroups] = {};
    auto flush_rows = [&](int g) -> int {
        if (!scoring[g]) return SDCA_OK;
        scoring[g] = false;
        hipEvent_t *ev = c->gev[g];
        HIP_TRY(hipEventSynchronize(ev[8]));
        c->timing.score += elapsed(ev[6], ev[7]);
        const float *h = c->h_rows + (size_t)g * gs * c->ws_L;
        for (size_t u = 0; u < row_items[g].size(); ++u) {
            // row k of the caller's buffer: at k * L (one alignment) or at row_off[k]
            const int item = row_items[g][u];
            const int La = c->members[alns ? alns[item] : 0]->L;
            const int64_t off = row_off ? row_off[item] : (int64_t)item * La;
            std::memcpy(score_rows + off, h + u * c->ws_L, (size_t)La * 4);
        }
        row_items[g].clear();
        return SDCA_OK;
    };
    int64_t rounds = 0;
    // the groups' rounds alternate: while the host steps one group's line searches, the
    // other group's round runs on the device
    while (done < K) {
        bool any = false;
        for (int g = 0; g < ng; ++g) {
            if (!inflight[g]) continue;
            any = true;
            inflight[g] = false;
            if ((rc = finish_round(c, g, S.cmds().data() + (size_t)g * gs))) return rc;
            ++rounds;
            finished.clear();
            for (int i = g * gs; i < (g + 1) * gs; ++i) {
                if (S.slot(i).phase == PH_IDLE) continue;
                if (S.process(i, c->h_res + (size_t)i * R.width())) finished.push_back(i);
            }
            if (!finished.empty()) {
                // score rows of the finished loci: the score kernel reads their points on
                // the group's stream ahead of the next round, which refills their slots,
                // so nothing here waits for the device (apart from x_out)
                hipStream_t st = c->gstream[g];
                hipEvent_t *ev = c->gev[g];
                const int nf = (int)finished.size();
                int64_t *h_off = c->h_src_off + (size_t)g * gs;
                int32_t *h_sl = c->h_score_loci + (size_t)g * gs;
                int32_t *h_ld = c->h_score_ld + (size_t)g * gs, *h_al = c->h_score_aln + (size_t)g * gs;
                for (int u = 0; u < nf; ++u) {
                    const int i = finished[u];
                    const Slot &sl = S.slot(i);
                    // the solution: the last trial point (packed rows of ws_ldx) or x_c (the
                    // slot vector, rows of its own Kxp)
                    h_off[u] = sl.from_xt ? c->h_xt_off[i] : c->xc_off + (int64_t)i * c->n_slot;
                    h_ld[u] = sl.from_xt ? (int32_t)c->ws_ldx : (int32_t)c->members[sl.aln]->Kxp;
                    h_sl[u] = sl.r;
                    h_al[u] = sl.aln;
                }
                int64_t *d_off = c->d_src_off + (size_t)g * gs;
                int32_t *d_sl = c->d_score_loci + (size_t)g * gs;
                int32_t *d_ld = c->d_score_ld + (size_t)g * gs, *d_al = c->d_score_aln + (size_t)g * gs;
                float *d_rows = c->d_rows + (size_t)g * gs * c->ws_L;
                HIP_TRY(hipMemcpyAsync(d_off, h_off, nf * 8, hipMemcpyHostToDevice, st));
                HIP_TRY(hipMemcpyAsync(d_sl, h_sl, nf * 4, hipMemcpyHostToDevice, st));
                HIP_TRY(hipMemcpyAsync(d_ld, h_ld, nf * 4, hipMemcpyHostToDevice, st));
                HIP_TRY(hipMemcpyAsync(d_al, h_al, nf * 4, hipMemcpyHostToDevice, st));
                HIP_TRY(hipEventRecord(ev[6], st));
                sdca::kern::score(st, c->d_Xtr, d_off, d_ld, d_sl, d_al, nf, c->ws_L, c->ws_L, c->d_aln, d_rows, 0);
                HIP_TRY(hipGetLastError());
                HIP_TRY(hipEventRecord(ev[7], st));
                if (score_rows) {
                    HIP_TRY(hipMemcpyAsync(c->h_rows + (size_t)g * gs * c->ws_L, d_rows, (size_t)nf * c->ws_L * 4,
                                           hipMemcpyDeviceToHost, st));
                    for (int u = 0; u < nf; ++u) row_items[g].push_back(S.slot(finished[u]).item);
                }
                HIP_TRY(hipEventRecord(ev[8], st));
                scoring[g] = true;
                if (x_out) {
                    for (int u = 0; u < nf; ++u) {
                        const Slot &sl = S.slot(finished[u]);
                        HIP_TRY(hipMemcpyAsync(xdev.data(), c->d_Xtr + h_off[u], c->n_slot * 4,
                                               hipMemcpyDeviceToHost, st));
                        HIP_TRY(hipStreamSynchronize(st));
                        dev_to_ref(c, sl.r, xdev.data(), x_out + (size_t)sl.item * n);
                    }
                }
                for (int u = 0; u < nf; ++u) {
                    const int i = finished[u];
                    const Slot &sl = S.slot(i);
                    if (st_out) {
                        sdca_locus_stats &o = st_out[sl.item];
                        o.fval = sl.f;
                        o.gnorm = sl.gnorm;
                        o.nfeval = sl.nfeval;
                        o.iter = sl.iter;
                        o.status = sl.status;
                        o.nfeval_first = sl.nfeval_first;
                        o.gnorm_J = sl.gnormJ;
                        o.nfeval_fine = sl.nfeval_fine;
                        o.nfeval_sub = 0;
                    }
                    noconv += (sl.status == SDCA_ST_ITER_CAP || sl.status == SDCA_ST_LS_FAIL);
                    nonfinite += (sl.status == SDCA_ST_NONFINITE);
                    ++done;
                    if ((rc = fill(i))) return rc;
                }
                if ((rc = stage_compact(g))) return rc;
                if (c->opts.verbose)
                    fprintf(stderr, "sdca: %d / %d loci solved after %lld rounds\n", done, K, (long long)rounds);
            }
            if (group_active(g)) {
                S.mark_live(g * gs, (g + 1) * gs);
                if ((rc = launch_round(c, g, S.cmds().data() + (size_t)g * gs))) return rc;
                inflight[g] = true;
            }
            if ((rc = flush_rows(g))) return rc;  // while the round runs
        }
        if (!any) break;  // every group idle
    }
    for (int g = 0; g < ng; ++g)
        if ((rc = flush_rows(g))) return rc;
    c->timing.total += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (nonfinite) return fail(SDCA_E_NONFINITE, "%d loci reached a non-finite objective", nonfinite);
    if (noconv) return fail(SDCA_E_NOCONV, "%d loci stopped without reaching the gradient threshold", noconv);
    return SDCA_OK;
}

extern "C" {

int sdca_pair_counts(sdca_ctx *c, const int32_t *pairs, int64_t npairs, int32_t *counts_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (npairs < 0 || (npairs > 0 && (!pairs || !counts_out))) return fail(SDCA_E_ARG, "bad arguments");
    for (int64_t p = 0; p < 2 * npairs; ++p)
        if (pairs[p] < 0 || pairs[p] >= c->L) return fail(SDCA_E_ARG, "pair locus %d out of range", pairs[p]);
    if (npairs == 0) return SDCA_OK;
    const int W = (c->B + 63) / 64;
    if (!c->d_SP) {
        if ((rc = dalloc(&c->d_SP, (size_t)c->L * c->q * W))) return rc;
        HIP_TRY(hipDeviceSynchronize());
        sdca::kern::seq_planes(c->stream, c->d_states, c->B, c->L, c->q, W, c->d_SP);
        HIP_TRY(hipGetLastError());
    }
    int32_t *d_pairs = nullptr, *d_counts = nullptr;
    const int64_t qq = (int64_t)c->q * c->q;
    if ((rc = dalloc(&d_pairs, (size_t)npairs * 2)) || (rc = dalloc(&d_counts, (size_t)npairs * qq))) {
        dfree(d_pairs);
        return rc;
    }
    hipError_t e = hipMemcpy(d_pairs, pairs, (size_t)npairs * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        sdca::kern::pair_counts(c->stream, c->d_SP, c->q, W, d_pairs, npairs, d_counts);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(counts_out, d_counts, (size_t)npairs * qq * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d_pairs);
    dfree(d_counts);
    if (e != hipSuccess) return fail(SDCA_E_HIP, "pair counts failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

int sdca_pair_count_table(sdca_ctx *c, int32_t i0, int32_t nloci, int32_t *out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (i0 < 0 || nloci < 1 || i0 + nloci > c->L || !out) return fail(SDCA_E_ARG, "bad row range");
    if ((rc = ensure_x8(c))) return rc;
    const int QP = c->QP;
    const int rows = nloci * QP;
    const int Mrows = (int)round_up(rows, sdca::kTile);
    int32_t *d_tab = nullptr;
    if ((rc = dalloc(&d_tab, (size_t)Mrows * c->Kc_pad))) return rc;
    HIP_TRY(hipDeviceSynchronize());
    sdca::kern::pair_gemm(c->stream, 0, QP, c->d_X8 + (size_t)i0 * QP * c->Bp, c->d_X8, c->Bp, Mrows, (int)c->Kc_pad,
                          i0, c->L, c->B, c->d_cnt, d_tab, nullptr);
    hipError_t e = hipGetLastError();
    const int64_t Kc = (int64_t)c->L * QP;
    if (e == hipSuccess)
        e = hipMemcpy2DAsync(out, Kc * 4, d_tab, c->Kc_pad * 4, Kc * 4, rows, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d_tab);
    if (e != hipSuccess) return fail(SDCA_E_HIP, "pair count table failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

int sdca_mi_topk(sdca_ctx *c, int32_t K, int32_t *partners_out, float *mi_out)
{
    int rc = check_ctx(c);
    if (rc) return rc;
    if (K < 1 || K > c->L - 1 || !partners_out) return fail(SDCA_E_ARG, "K must be in [1, L-1] (got %d)", K);
    if ((rc = ensure_x8(c))) return rc;
    const int QP = c->QP, L = c->L;
    const int Mrows = (int)std::min<int64_t>(1024, round_up((int64_t)L * QP, sdca::kTile));
    const int PL = Mrows / QP;  // loci per panel
    float *d_mi = nullptr, *d_val = nullptr;
    int32_t *d_idx = nullptr;
    auto cleanup = [&]() {
        dfree(d_mi);
        dfree(d_val);
        dfree(d_idx);
    };
    if ((rc = dalloc(&d_mi, (size_t)PL * L)) || (rc = dalloc(&d_idx, (size_t)L * K)) ||
        (mi_out && (rc = dalloc(&d_val, (size_t)L * K)))) {
        cleanup();
        return rc;
    }
    HIP_TRY(hipDeviceSynchronize());
    hipError_t e = hipSuccess;
    for (int p0 = 0; p0 < L && e == hipSuccess; p0 += PL) {
        sdca::kern::pair_gemm(c->stream, 1, QP, c->d_X8 + (size_t)p0 * QP * c->Bp, c->d_X8, c->Bp, Mrows,
                              (int)c->Kc_pad, p0, L, c->B, c->d_cnt, nullptr, d_mi);
        sdca::kern::topk(c->stream, d_mi, std::min(PL, L - p0), L, K, p0, d_idx, d_val);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(partners_out, d_idx, (size_t)L * K * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess && mi_out)
        e = hipMemcpyAsync(mi_out, d_val, (size_t)L * K * 4, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {  // keep the partner table for warm starts
        dfree(c->d_topk);
        c->d_topk = d_idx;
        c->topk_K = K;
        d_idx = nullptr;
    }
    cleanup();
    if (e != hipSuccess) return fail(SDCA_E_HIP, "MI top-K failed: %s", hipGetErrorString(e));
    return SDCA_OK;
}

int sdca_gemm_config(const sdca_ctx *c, int32_t cfg[7])
{
    if (!c || !cfg) return fail(SDCA_E_ARG, "null argument");
    if (!c->ws_slots) return fail(SDCA_E_STATE, "no workspace yet");
    cfg[0] = c->tuned;
    cfg[1] = c->ksplit;
    cfg[2] = c->kchunk;
    cfg[3] = c->gm_energy;
    cfg[4] = c->gm_grad;
    cfg[5] = c->tn_energy;
    cfg[6] = c->tn_grad;
    return SDCA_OK;
}

int sdca_get_timing(sdca_ctx *c, sdca_timing *t, int32_t reset)
{
    if (!c || !t) return fail(SDCA_E_ARG, "null argument");
    const auto &s = c->timing;
    t->gemm_energy_ms = s.energy;
    t->gemm_grad_ms = s.grad;
    t->softmax_ms = s.softmax;
    t->lbfgs_ms = s.lbfgs;
    t->score_ms = s.score;
    t->reweight_ms = s.reweight;
    t->total_ms = s.total;
    t->n_evals = s.evals;
    t->n_energy_launches = s.energy_launches;
    t->n_grad_launches = s.grad_launches;
    t->flops_energy = s.flops_energy;
    t->flops_grad = s.flops_grad;
    t->bytes_lbfgs = s.bytes_lbfgs;
    t->update_ms = s.update;
    t->dots_ms = s.dots;
    t->bytes_update = s.bytes_update;
    t->bytes_dots = s.bytes_dots;
    t->subspace_ms = s.sub;
    t->n_sub_evals = s.sub_evals;
    t->flops_sub = s.flops_sub;
    t->host_ms = s.host;
    t->gstep_ms = s.gstep;
    t->bytes_gstep = s.bytes_gstep;
    if (reset) c->timing = sdca::Timing();
    return SDCA_OK;
}

}  // extern "C"

// ====================================================== multi-alignment solves

// A solve over several alignments at once: an internal context that holds the workspace,
// streams and events, whose members are the caller's contexts (their one-hot X, weights
// and lambda); the caller's contexts must outlive it.
struct sdca_multi {
    sdca_ctx *w = nullptr;
    ~sdca_multi() { delete w; }
};

extern "C" {

int sdca_multi_create(sdca_multi **out, sdca_ctx *const *ctxs, int32_t n, const sdca_opts *opts)
{
    if (!out) return fail(SDCA_E_ARG, "out is null");
    *out = nullptr;
    if (!ctxs || n < 1 || n > 4096) return fail(SDCA_E_ARG, "need 1 to 4096 contexts (got %d)", n);
    sdca_opts o;
    if (opts) o = *opts;
    else sdca_default_opts(&o);
    if (o.precision != SDCA_PREC_FP32 && o.precision != SDCA_PREC_SPLIT_BF16)
        return fail(SDCA_E_ARG, "unknown precision %d", o.precision);
    if (o.formulation < SDCA_FORM_AUTO || o.formulation > SDCA_FORM_GRAM)
        return fail(SDCA_E_ARG, "unknown formulation %d", o.formulation);
    if (!valid_history(o.history)) return fail(SDCA_E_ARG, "history must be 0 (the formulation's default) or one of 1, 3, 5, 8, 10, 16, 24, 32 (got %d)", o.history);
    if (o.batch_loci < 0 || o.max_iter < 1 || o.max_ls < 1 || !(o.grad_threshold >= 0.0) || !(o.j_threshold >= 0.0) ||
        o.coarse_factor < 0 || o.fine_factor < 0 || (o.subspace_rank != 0 && o.subspace_rank != sdca::kSubRank))
        return fail(SDCA_E_ARG, "invalid solver options");
    for (int a = 0; a < n; ++a) {
        if (!ctxs[a]) return fail(SDCA_E_ARG, "context %d is null", a);
        if (ctxs[a]->device != o.device)
            return fail(SDCA_E_ARG, "context %d is on device %d, the solve on %d", a, ctxs[a]->device, o.device);
    }
    auto *m = new sdca_multi();
    m->w = new sdca_ctx();
    sdca_ctx *w = m->w;
    w->opts = o;
    w->opts.pipeline_groups = 1;
    w->device = o.device;
    int rc = create_streams(w);
    if (rc) {
        delete m;
        return rc;
    }
    w->members.assign(ctxs, ctxs + n);
    *out = m;
    return SDCA_OK;
}

void sdca_multi_destroy(sdca_multi *m) { delete m; }

int sdca_multi_solve(sdca_multi *m, const int32_t *alns, const int32_t *loci, int32_t K, float *score_rows,
                     const int64_t *row_off, sdca_locus_stats *st)
{
    if (!m || !m->w) return fail(SDCA_E_ARG, "null multi-alignment solve");
    if (K > 0 && !alns) return fail(SDCA_E_ARG, "alns is null");
    int rc = check_ctx(m->w);
    if (rc) return rc;
    std::vector<int64_t> packed;
    if (score_rows && !row_off) {  // rows one after the other, each the L of its alignment
        packed.resize(K);
        int64_t off = 0;
        for (int k = 0; k < K; ++k) {
            if (alns[k] < 0 || alns[k] >= (int)m->w->members.size())
                return fail(SDCA_E_ARG, "alignment %d of item %d out of range", alns[k], k);
            packed[k] = off;
            off += m->w->members[alns[k]]->L;
        }
        row_off = packed.data();
    }
    return sdca_solve_batch(m->w, alns, loci, K, nullptr, nullptr, score_rows, row_off, nullptr, st);
}

int sdca_multi_get_timing(sdca_multi *m, sdca_timing *t, int32_t reset)
{
    if (!m) return fail(SDCA_E_ARG, "null multi-alignment solve");
    return sdca_get_timing(m->w, t, reset);
}

}  // extern "C"

