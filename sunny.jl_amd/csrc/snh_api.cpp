// C-ABI of libsunnyhip: device-resident System (and replica batches), the
// step!/energy entry points, and their orchestration on one HIP stream.
#include "../../include/sunny_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cstdlib>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "snh_internal.h"

using namespace snh;

namespace {

thread_local std::string g_create_error;

struct Failure {
    int code;
    std::string msg;
};

#define SNH_HIP(call)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw Failure{SNH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)};             \
    } while (0)

#define SNH_REQUIRE(cond, code, msg)                                                                    \
    do {                                                                                                \
        if (!(cond)) throw Failure{code, msg};                                                          \
    } while (0)

template <typename X>
X* dalloc(size_t n)
{
    void* p = nullptr;
    if (n == 0) n = 1;
    SNH_HIP(hipMalloc(&p, n * sizeof(X)));
    return static_cast<X*>(p);
}

template <typename X>
X* dupload(const std::vector<X>& v)
{
    X* p = dalloc<X>(v.size());
    if (!v.empty()) SNH_HIP(hipMemcpy(p, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice));
    return p;
}

// Stevens quadrupoles O[2,q], q = 2..-2, in the N-dim spin irrep, as
// polynomials of spin matrices (src/Operators/Stevens.jl:12-86 at k = 2).
std::vector<double> quadrupole_ops(int N)
{
    using cd = std::complex<double>;
    double s = (N - 1) / 2.0;
    std::vector<cd> Sx(N * N), Sy(N * N), Sz(N * N);
    for (int j = 1; j < N; j++) {
        double off = std::sqrt(2 * (s + 1) * j - (double)j * (j + 1)) / 2;
        Sx[(j - 1) * N + j] = off;
        Sx[j * N + j - 1] = off;
        Sy[(j - 1) * N + j] = cd(0, -off);
        Sy[j * N + j - 1] = cd(0, off);
    }
    for (int k = 0; k < N; k++) Sz[k * N + k] = s - k;
    auto mul = [N](const std::vector<cd>& A, const std::vector<cd>& B) {
        std::vector<cd> C(N * N);
        for (int a = 0; a < N; a++)
            for (int b = 0; b < N; b++) {
                cd acc = 0;
                for (int k = 0; k < N; k++) acc += A[a * N + k] * B[k * N + b];
                C[a * N + b] = acc;
            }
        return C;
    };
    auto xx = mul(Sx, Sx), yy = mul(Sy, Sy), zz = mul(Sz, Sz), xz = mul(Sx, Sz), zx = mul(Sz, Sx),
         yz = mul(Sy, Sz), zy = mul(Sz, Sy), xy = mul(Sx, Sy), yx = mul(Sy, Sx);
    std::vector<double> out(5 * N * N * 2);
    for (int k = 0; k < N * N; k++) {
        cd q[5] = {xx[k] - yy[k], (xz[k] + zx[k]) / 2.0, 2.0 * zz[k] - xx[k] - yy[k], (yz[k] + zy[k]) / 2.0,
                   xy[k] + yx[k]};
        for (int p = 0; p < 5; p++) {
            out[2 * (p * N * N + k)] = q[p].real();
            out[2 * (p * N * N + k) + 1] = q[p].imag();
        }
    }
    return out;
}

// LDS tile plan for the :dipole gather: tile (T1, T2), workgroup size and
// the split of axis 3 into segments, chosen by a cost model over all
// candidates that fit the LDS.  Cost of one workgroup = (layers loaded) x
// (positions per layer) + 3 x (site lanes computed, idle lanes included),
// each divided by the fraction of a 128-byte line its rows fill; a CU works
// through ceil(workgroups / CUs) of them, 1.3x slower when only one fits.
struct TilePlan {
    TileArgs A{};
    int threads = 0;
    size_t lds = 0;
    std::vector<TileWork> work;
    std::vector<int> rwork;  // [R+1] workgroup range per replica
};

int env_int(const char* name, int dflt)
{
    const char* v = std::getenv(name);
    return (v && *v) ? std::atoi(v) : dflt;
}

// Work list of a plan whose A.nwork holds its segment count along axis 3.
void tile_work_list(TilePlan& P, const std::vector<std::array<int, 3>>& dims)
{
    const int nseg = P.A.nwork;
    P.work.clear();
    P.rwork.assign(dims.size() + 1, 0);
    for (size_t r = 0; r < dims.size(); r++) {
        P.rwork[r] = (int)P.work.size();
        const int L3 = dims[r][2];
        const int ns = std::min(nseg, L3);
        for (int sgm = 0; sgm < ns; sgm++) {
            const int k0 = (int)((long long)L3 * sgm / ns), k1 = (int)((long long)L3 * (sgm + 1) / ns);
            for (int t2 = 0; t2 < dims[r][1]; t2 += P.A.T2)
                for (int t1 = 0; t1 < dims[r][0]; t1 += P.A.T1) P.work.push_back({(int)r, t1, t2, k0, k1, {0, 0, 0}});
        }
    }
    P.rwork[dims.size()] = (int)P.work.size();
    P.A.nwork = (int)P.work.size();
}

// Tile plans of the cost model, best first (at most `keep`, work lists built).
bool plan_tiles(int na, const int w[3], const std::vector<std::array<int, 3>>& dims, size_t elem, size_t lds_wg,
                size_t lds_cu, int ncu, std::vector<TilePlan>& ranked, size_t keep)
{
    std::vector<std::pair<double, TilePlan>> cands;
    auto pow2 = [](int v) {
        int p = 1;
        while (p < v) p <<= 1;
        return p;
    };
    int L1max = 1, L2max = 1, L3min = INT_MAX;
    for (const auto& d : dims) {
        L1max = std::max(L1max, d[0]);
        L2max = std::max(L2max, d[1]);
        L3min = std::min(L3min, d[2]);
    }
    const int nslot = 2 * w[2] + 1;
    // overrides for tuning and tests: SNH_TILE_T1, SNH_TILE_T2, SNH_TILE_NSEG
    const int want_T1 = env_int("SNH_TILE_T1", 0), want_T2 = env_int("SNH_TILE_T2", 0),
              want_seg = env_int("SNH_TILE_NSEG", 0), want_sp = env_int("SNH_TILE_SP", 0);
    for (int sp = 1; sp <= 2; sp++)
    for (int T1 = std::min(64, pow2(L1max)); T1 >= 1; T1 >>= 1)
        for (int T2 = 1; T2 <= std::min(64, pow2(L2max)); T2 <<= 1) {
            if ((want_T1 && T1 != want_T1) || (want_T2 && T2 != want_T2) || (want_sp && sp != want_sp)) continue;
            // workgroup size limits of the kernels (tile_threads_max)
            const int max_threads = (elem == 4 && sp == 1) ? 1024 : 512;
            const int sites_per_thread = sp;
            const long long ROW = T1 + 2 * w[0], ROWS = T2 + 2 * w[1];
            const long long npos = ROWS * na * ROW;
            const long long spl = (T1 * T2 + 63) / 64 * 64;
            const long long nsite = na * spl;
            const int load_per = tile_load_per(sp);
            long long thr = std::max((nsite + sites_per_thread - 1) / sites_per_thread,
                                     (npos + load_per - 1) / load_per);
            thr = (thr + 63) / 64 * 64;
            if (thr > max_threads) continue;
            const size_t bytes = (size_t)nslot * npos * 3 * elem;
            if (bytes > lds_wg) continue;
            // workgroups resident per CU: LDS, 2048 threads, and the VGPR
            // budget (about 88 / 128 VGPRs for FP32 with 1 / 2 sites per
            // thread, 150 for FP64; 512 VGPRs per SIMD lane, 4 SIMDs)
            const int vgpr = elem == 4 ? (sp == 1 ? 88 : 128) : 150;
            const long long waves_cu = 4LL * (512 / vgpr), waves_wg = thr / 64;
            const int per_cu = (int)std::max<long long>(
                1, std::min({(long long)(lds_cu / bytes), 2048 / thr, waves_cu / waves_wg}));
            // below 16 resident waves per CU the sweep cannot hide its load
            // latency; the ghost-padded gather then does better (an explicit
            // tile request is honoured regardless)
            const bool forced = want_T1 || want_T2 || want_seg || want_sp;
            if (!forced && per_cu * waves_wg < 16) continue;
            long long base = 0;
            for (const auto& d : dims) base += (long long)((d[0] + T1 - 1) / T1) * ((d[1] + T2 - 1) / T2);
            for (int nseg = want_seg ? std::min(want_seg, L3min) : 1; nseg <= L3min; nseg *= 2) {
                if (want_seg && nseg != std::min(want_seg, L3min)) break;
                const long long nwork = base * nseg;
                // a CU retires workgroups one after another; a lone workgroup
                // per CU idles at its barriers (factor 1.3)
                const long long rounds = (nwork + ncu - 1) / ncu;
                const long long kseg = (L3min + nseg - 1) / nseg;
                // rows shorter than a 128-byte line waste load and store bandwidth
                const double eff_ld = std::min(1.0, (double)ROW * elem / 128.0);
                const double eff_st = std::min(1.0, (double)T1 * elem / 128.0);
                const double bubble = per_cu >= 2 ? 1.0 : 1.3;
                const double cost = (double)rounds * bubble *
                                    ((double)(kseg + 2 * w[2]) * npos / eff_ld + 3.0 * kseg * nsite / eff_st);
                {
                    TilePlan P;
                    TileArgs& A = P.A;
                    A.T1 = T1;
                    A.T2 = T2;
                    A.lt1 = 0;
                    while ((1 << A.lt1) < T1) A.lt1++;
                    A.w1 = w[0];
                    A.w2 = w[1];
                    A.w3 = w[2];
                    A.ROW = (int)ROW;
                    A.nslot = nslot;
                    A.npos = (int)npos;
                    A.cap = (int)npos;
                    A.spl = (int)spl;
                    A.nsite = (int)nsite;
                    A.sp = sp;
                    P.threads = (int)thr;
                    P.lds = bytes;
                    A.nwork = nseg;  // segments per column until the work list is built
                    cands.push_back({cost, std::move(P)});
                }
            }
        }
    if (cands.empty()) return false;
    std::stable_sort(cands.begin(), cands.end(),
                     [](const std::pair<double, TilePlan>& a, const std::pair<double, TilePlan>& b) {
                         return a.first < b.first;
                     });
    ranked.clear();
    for (size_t c = 0; c < cands.size() && c < keep; c++) {
        ranked.push_back(std::move(cands[c].second));
        tile_work_list(ranked.back(), dims);
    }
    return true;
}

// LDS byte offset of every (padded) bond from its site, per ring phase p:
// neighbour (j, c + n) of a site (i, c) in layer k sits n1 + ROW ((j - i) +
// na n2) positions away within a slot, in the slot of layer k + n3, which is
// ((p + n3) mod nslot - p) slots away.  bsrc[k] < 0 marks padding (offset 0).
std::vector<int> tile_lds_offsets(const TileArgs& A, int na, const int32_t* bond, const std::vector<int>& bsrc,
                                  size_t elem)
{
    const size_t nbp = bsrc.size();
    std::vector<int> off((size_t)A.nslot * nbp, 0);
    for (int ph = 0; ph < A.nslot; ph++)
        for (size_t k = 0; k < nbp; k++) {
            if (bsrc[k] < 0) continue;
            const int32_t* bd = bond + 6 * bsrc[k];
            const int dslot = ((ph + bd[4]) % A.nslot + A.nslot) % A.nslot - ph;
            const long long pos =
                bd[2] + (long long)A.ROW * ((bd[1] - bd[0]) + (long long)na * bd[3]) + (long long)dslot * A.cap;
            off[(size_t)ph * nbp + k] = (int)(pos * 3 * (long long)elem);
        }
    return off;
}

struct KernelTimer {
    bool enabled = false;
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> events;
    std::vector<hipEvent_t> pool;

    hipEvent_t get()
    {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        SNH_HIP(hipEventCreate(&e));
        return e;
    }
    void clear()
    {
        for (auto& p : events) {
            pool.push_back(p.second.first);
            pool.push_back(p.second.second);
        }
        events.clear();
    }
    ~KernelTimer()
    {
        clear();
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Colour classes for parallel Metropolis sweeps.  Class (i, r) holds the
// sites (c, i) with c = r mod m (componentwise); m divides every replica's
// dims, so a bond (i -> j, n) joins class (i, r) only to class (j, r + n mod m).
// A proper colouring of that quotient graph leaves no bond inside a colour.
// Tries periods m up to 8 along each axis (and the full common period) and
// keeps the colouring with the fewest colours (then the smallest period).
struct Coloring {
    int m[3] = {1, 1, 1};
    int ncolors = 0;
    std::vector<int> color;  // [na][m3][m2][m1]
};

Coloring color_classes(int na, int nbonds, const int32_t* bond, const std::vector<std::array<int, 3>>& dims)
{
    int g[3] = {0, 0, 0};
    for (const auto& d : dims)
        for (int a = 0; a < 3; a++) g[a] = std::gcd(g[a], d[a]);
    std::vector<int> cand[3];
    for (int a = 0; a < 3; a++)
        for (int m = 1; m <= g[a]; m++)
            if (g[a] % m == 0 && (m <= 8 || m == g[a])) cand[a].push_back(m);
    Coloring best;
    long long best_nodes = 0;
    for (int m1 : cand[0])
        for (int m2 : cand[1])
            for (int m3 : cand[2]) {
                const long long nodes = (long long)na * m1 * m2 * m3;
                if (nodes > (1 << 20)) continue;
                const int m[3] = {m1, m2, m3};
                auto node = [&](int i, const int r[3]) { return r[0] + m1 * (r[1] + m2 * (r[2] + m3 * i)); };
                std::vector<int> color(nodes, -1);
                bool loop = false;
                int nc = 0;
                std::vector<char> used;
                for (long long u = 0; u < nodes && !loop; u++) {
                    const int i = (int)(u / ((long long)m1 * m2 * m3));
                    long long t = u % ((long long)m1 * m2 * m3);
                    const int r[3] = {(int)(t % m1), (int)((t / m1) % m2), (int)(t / ((long long)m1 * m2))};
                    used.assign(nc + 1, 0);
                    for (int b = 0; b < nbonds; b++) {
                        const int32_t* bd = bond + 6 * b;
                        if (bd[0] != i) continue;
                        int rr[3];
                        for (int a = 0; a < 3; a++) rr[a] = (((r[a] + bd[2 + a]) % m[a]) + m[a]) % m[a];
                        const int v = node(bd[1], rr);
                        if (v == u) {
                            loop = true;
                            break;
                        }
                        if (color[v] >= 0) used[color[v]] = 1;
                    }
                    int c = 0;
                    while (c < nc && used[c]) c++;
                    color[u] = c;
                    nc = std::max(nc, c + 1);
                }
                if (loop) continue;
                if (best.ncolors == 0 || nc < best.ncolors || (nc == best.ncolors && nodes < best_nodes)) {
                    best.m[0] = m1;
                    best.m[1] = m2;
                    best.m[2] = m3;
                    best.ncolors = nc;
                    best.color = color;
                    best_nodes = nodes;
                }
            }
    return best;
}

}  // namespace

// ---------------------------------------------------------------------------

struct snh_system {
    std::string err;
    virtual ~snh_system() = default;
    virtual int64_t total_sites() const = 0;
    virtual int replicas() const = 0;
    virtual int64_t replica_sites(int r) const = 0;
    virtual void set_state(const double* x, bool coherent) = 0;
    virtual void get_state(double* x, bool coherent) = 0;
    virtual void step_langevin(double dt, double damping, const double* kT, const double* noise, uint64_t seed,
                               uint64_t step0, int nsteps) = 0;
    virtual void step_midpoint(double dt, double damping, double tol, int max_iters, int nsteps, int* iters) = 0;
    virtual void energy(double* E) = 0;
    virtual void energy_noalloc(double* E_dev) = 0;
    virtual void energy_grad(double* out) = 0;
    virtual void synchronize() = 0;
    virtual void enable_timing(bool on) = 0;
    virtual void kernel_time(const char* fam, double* ms, int64_t* n) = 0;
    virtual void copy_calibration(int repeats, int64_t* bytes) = 0;
    virtual int64_t query(const std::string& key) const = 0;
    virtual void autotune(int candidates, int launches, double* best_ms) = 0;
    virtual void corr_init(int nobs, const double* obs, int ncorr, const int32_t* pairs, int T, int measperiod,
                           int window) = 0;
    virtual void corr_add_sample(double dt, double tol, int max_iters) = 0;
    virtual void corr_data(double* out, int64_t* nsamples) = 0;
    virtual void step_local(const double* kT, int nprop, const int32_t* kinds, const double* mags,
                            const double* weights, uint64_t seed, uint64_t sweep0, double nsweeps,
                            double* dE_dS) = 0;
    virtual void local_coloring(int32_t* m, int32_t* ncolors, int32_t* color) = 0;
    virtual void midpoint_record(double dt, double tol, int max_iters, int nsteps, int every) = 0;
    virtual void midpoint_adjoint(int nparams, const snh_spec* params, const double* dLdS, double* grad,
                                  double* dLdS0) = 0;
};

namespace {

template <typename T>
struct SystemImpl final : snh_system {
    hipStream_t st = nullptr;
    int dev = 0;  // the device the handle was created on
    int R = 0, na = 0, N = 0, nb = 0, nblocks = 0;
    uint32_t total = 0;  // padded sites over all replicas (plane stride)
    Ghost G{};
    bool ewald = false;
    std::vector<ReplicaInfo> reps;
    std::vector<int> rblk;
    // LDS-tiled gather (unpadded layout) when it fits; else ghost-padded gather
    bool tiled = false;
    // Tiled plans: tp for the predictor, energy and gradient; tc for the
    // corrector and the midpoint iteration (the kernels share the state
    // layout, so each may run on its own plan).
    TilePlan tp, tc;
    std::vector<TilePlan> tile_cands;  // the cost model's best plans, best first
    std::vector<void*> plan_mem[2];    // device tables of tp and tc
    int* d_tcblk = nullptr;            // per-replica workgroup ranges of tc
    size_t partial_cap = 0;            // entries of `partial`
    int tuned_index[2] = {-1, -1};     // candidates kept by autotune (-1: the cost model's first)
    int* d_trblk = nullptr;
    // bond tables as given (for the adjoint's per-parameter unit couplings)
    std::vector<int> bsrc_h;      // padded slot -> original bond, -1 for padding
    std::vector<int> bflag_h;     // padded slot flags (kBond*)
    std::vector<int32_t> bond_h;  // [nb_in][6]
    int nb_in = 0;
    std::vector<int> kmax_h;
    bool any_biquad = false;
    // recorded midpoint trajectory (checkpoints on the device)
    std::vector<T*> ckpt;
    std::vector<int> rec_iters;
    int rec_every = 0, rec_n = 0, rec_maxit = 0;
    double rec_dt = 0, rec_tol = 0;
    // device tables
    std::vector<void*> owned;
    Couplings<T> C{};
    int* d_rblk = nullptr;
    // state
    T* X = nullptr;     // current state (3 or 2N planes)
    T* Xp = nullptr;    // Heun stage S' (with ghosts)
    T* Xo = nullptr;    // Heun half-update H (own sites only); scratch planes
    T* B1 = nullptr;    // midpoint buffers (lazy)
    T* B2 = nullptr;
    T* Bh = nullptr;    // materialised midpoint state: Shat (:dipole) or Zbar (SU(N) + Ewald)
    T* Bh2 = nullptr;   // :dipole: the next iteration's Shat
    T* Bmu = nullptr;   // tiled :dipole: mu = |(S + S'')/2| / kappa of odd ([0]) and even ([total]) iterations
    T* Sd = nullptr;    // expected dipoles of a coherent state (SU(N) + Ewald)
    T* Dz = nullptr;    // SU(N) Langevin without biquadratic bonds: expected dipoles
    T* Dzp = nullptr;   //   of the state and of the predictor stage, with images
    bool dz_valid = false;  // Dz holds the dipoles of X as the last corrector wrote them
    double* partial = nullptr;
    double* d_out = nullptr;   // [R]
    int* conv_iter = nullptr;  // [R]
    double* atol2 = nullptr;   // [R]
    T* nscale = nullptr;       // [R]
    double* host_noise = nullptr;
    double* pinned = nullptr;  // [R] host readback
    int* pinned_i = nullptr;   // [R]
    int planes = 3;
    int last_iters = 4;
    // Small :dipole replicas on the ghost layout (no Ewald) step in one
    // persistent workgroup per replica: no launch per stage, no host check
    bool coop = false;  // cooperative midpoint launch (one replica, ghost layout)
    double* coop_partial = nullptr;  // [2 nblocks]
    unsigned* coop_bar = nullptr;    // count, gen, err
    // per midpoint step.
    bool persist = false;
    // Tiled Langevin batches of two or more replicas (no Ewald) step as two
    // replica groups on two streams in anti-phase: one group's corrector
    // (latency-bound) runs beside the other's predictor (VALU-bound).
    int pair_split = 0;        // first replica of the second group (0: no groups)
    bool pair_on = false;      // step in groups: SNH_PAIR=2, or chosen by snh_autotune (SNH_PAIR=1)
    bool mid_sync = false;     // SNH_MID_SYNC=1: host convergence check after every midpoint step
    bool mid_hat_ready = false;  // :dipole: the last finalize wrote the next step's first Shat (into Bh)
    int* last_it = nullptr;    // [R] iterations of the last step of a budgeted group
    int64_t mid_reruns = 0;     // midpoint step groups undone after overrunning their budget
    int64_t mid_fallbacks = 0;  // midpoint step groups then run with per-step host checks
    hipStream_t st2 = nullptr;
    hipEvent_t pev[3] = {nullptr, nullptr, nullptr};  // start, first group's predictor, second group's predictor
    int* fail_step = nullptr;  // [R] persistent midpoint / budgeted groups: step that did not converge
    int* max_iter = nullptr;   // [R] budgeted midpoint groups: largest iteration count
    // Sampled correlations of one :dipole replica (snh_corr.hip)
    struct Corr {
        int nobs = 0, ncorr = 0, nt = 0, measperiod = 0;  // nt: time slots T = n_all_omega
        long long ncell = 0;
        int L[3] = {1, 1, 1};
        std::vector<int> pairs;  // [ncorr][2] observable indices (alpha, beta)
        std::vector<void*> mem;
        double* obs = nullptr;   // [nobs][na][3]
        void* samp = nullptr;    // [T][nobs][na][ncell] complex: A(r, t), then A(q, t)
        void* samp2 = nullptr;   // [T][nobs][na][ncell] complex: A(q, w)
        void* P = nullptr;       // [T][na][ncell] one channel's products
        void* Y = nullptr;       // [T][na][ncell] its windowed real-time correlation
        void* data = nullptr;    // [ncorr][na][T][na][ncell] running mean
        double* post = nullptr;  // [T] window / (T n_contrib)
        long long nsamples = 0;
        CorrLibFft* lib = nullptr;  // hipFFT plan of the lattice transform (axes beyond 32)
    } corr;
    // Metropolis sweeps (snh_mc.hip): colour classes and tables, built at first use
    struct Mc {
        bool ready = false;
        Coloring col;
        bool wide = false;            // some |n| exceeds a replica's dims
        std::vector<int> coff;        // [ncolors + 1] work items of each colour
        int* bn = nullptr;            // [nb][4]
        McWork* work = nullptr;       // colour-major, then replica, then class
        int* start = nullptr;         // [ncolors][R + 1] work ranges of each replica
        double* kT = nullptr;         // [R]
        double* acc = nullptr;        // [max(work, R)][4]
        double* out = nullptr;        // [R][4]
        // sequential trials with the Ewald term (tensors per Ewald group)
        double* phi[2] = {nullptr, nullptr};    // FP64 potential planes [3][total], alternating
        // accepted moves of a chunk, [kMcFftPeriod][R][kMcChunk], and their
        // counts [kMcFftPeriod][R]: slots 0 and 1 alternate for the apply
        // path, FFT groups fill slot k mod kMcFftPeriod between refreshes
        McTrial* moves = nullptr;
        int* nmoves = nullptr;
        double* pprev = nullptr;  // [R][kMcChunk][3] earlier moves at the next chunk's trial sites
    } mc;
    // Ewald: one group per run of consecutive replicas with equal dims, each
    // with its grid, transformed tensor, spectra and (library FFT path) plans
    struct EwaldGroup {
        int r0 = 0, nr = 0;  // replicas r0 .. r0 + nr - 1
        EwaldGeom EG{};
        EwaldAParams P{};  // parameters of the real-space tensor
        void* M = nullptr;
        void* Fmu = nullptr;
        void* Fphi = nullptr;
        EwaldLibFft* lib = nullptr;
        double* Aew = nullptr;  // LocalSampler: the real-space tensor, kept (six doubles per block)
        double* A0 = nullptr;   // LocalSampler: [na][9] self blocks A(0, i, i)
    };
    std::vector<EwaldGroup> ewg;
    std::vector<double> ew_pos;  // fractional positions [na][3]
    T* phi1 = nullptr;
    T* phi2 = nullptr;
    KernelTimer timer;
    // Captured Langevin steps: a chunk of graph_steps steps with the noise
    // step counter in device memory, replayed while the launch parameters
    // (dt, damping, seed, noise mode, neighbour source) stay those of `gkey`.
    unsigned long long* d_step = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    std::array<double, 5> gkey{};
    int graph_steps = 0;
    bool graph_ok = true;
    bool warmed = false;

    template <typename X_>
    X_* own(X_* p)
    {
        owned.push_back(p);
        return p;
    }

    ~SystemImpl() override
    {
        if (st) (void)hipStreamSynchronize(st);
        if (st2) (void)hipStreamSynchronize(st2);
        for (hipEvent_t e : pev)
            if (e) (void)hipEventDestroy(e);
        if (st2) (void)hipStreamDestroy(st2);
        drop_graph();
        for (void* p : corr.mem) (void)hipFree(p);
        corr_lib_destroy(corr.lib);
        timer.clear();
        for (T* p : ckpt) (void)hipFree(p);
        for (auto& g : ewg) ewald_lib_destroy(g.lib);
        for (void* p : owned) (void)hipFree(p);
        if (pinned) (void)hipHostFree(pinned);
        if (pinned_i) (void)hipHostFree(pinned_i);
        if (st) (void)hipStreamDestroy(st);
    }

    void build(const snh_spec& sp)
    {
        // graph replays measured slower than eager launches on ROCm 7.2 (about
        // 3.7 us more per kernel node): off unless SNH_GRAPH_STEPS asks
        graph_steps = std::max(0, env_int("SNH_GRAPH_STEPS", 0));
        R = sp.nreplicas;
        na = sp.na;
        N = sp.N;
        nb = sp.nbonds;
        ewald = sp.ewald != 0;
        planes = (N == 0) ? 3 : 2 * N;
        SNH_HIP(hipGetDevice(&dev));
        SNH_HIP(hipStreamCreate(&st));
        // this device's copy of the static scratch of snh_energy_noalloc
        SNH_HIP(noalloc_scratch_init());

        // ghost widths: the longest bond offset along each axis
        int w[3] = {0, 0, 0};
        for (int b = 0; b < nb; b++) {
            const int* bd = sp.bond + 6 * b;
            SNH_REQUIRE(bd[0] >= 0 && bd[0] < na && bd[1] >= 0 && bd[1] < na, SNH_ERR_INVALID,
                        "bond atom index out of range");
            for (int a = 0; a < 3; a++) w[a] = std::max(w[a], std::abs(bd[2 + a]));
        }
        SNH_REQUIRE(sp.layout == SNH_LAYOUT_AUTO || sp.layout == SNH_LAYOUT_GHOST, SNH_ERR_INVALID,
                    "layout must be SNH_LAYOUT_AUTO or SNH_LAYOUT_GHOST");
        // one small :dipole replica without Ewald: midpoint steps in one
        // cooperative launch (ghost layout), above the one-workgroup size
        {
            const long long sites = sp.nreplicas == 1 ? (long long)sp.dims[0] * sp.dims[1] * sp.dims[2] * na : 0;
            coop = N == 0 && !ewald && R == 1 && sites > kPersist &&
                   sites <= env_int("SNH_COOP_MAX_SITES", kCoopMaxSites) && env_int("SNH_COOP", 1) != 0;
        }
        if (N == 0 && sp.layout == SNH_LAYOUT_AUTO && !coop) {
            int dev = 0;
            SNH_HIP(hipGetDevice(&dev));
            hipDeviceProp_t prop;
            SNH_HIP(hipGetDeviceProperties(&prop, dev));
            std::vector<std::array<int, 3>> dv(R);
            for (int r = 0; r < R; r++) {
                dv[r] = {sp.dims[3 * r], sp.dims[3 * r + 1], sp.dims[3 * r + 2]};
                for (int a = 0; a < 3; a++) SNH_REQUIRE(dv[r][a] >= 1, SNH_ERR_INVALID, "replica dims must be positive");
            }
            const size_t lds_cu = prop.maxSharedMemoryPerMultiProcessor ? prop.maxSharedMemoryPerMultiProcessor
                                                                         : prop.sharedMemPerBlock;
            // 256 bytes stay free for the kernels' static LDS (block reductions)
            tiled = plan_tiles(na, w, dv, sizeof(T), prop.sharedMemPerBlock - 256, lds_cu,
                               std::max(1, prop.multiProcessorCount), tile_cands, 16);
            if (tiled) tp = tile_cands[0];
        }
        // the tiled sweep fetches halos with the periodic wrap itself: no ghosts
        const int lw[3] = {tiled ? 0 : w[0], tiled ? 0 : w[1], tiled ? 0 : w[2]};
        G.w1 = lw[0];
        G.w2 = lw[1];
        G.w3 = lw[2];
        G.general = 0;

        // replicas and dims classes
        std::map<std::vector<int>, int> dcls_of;
        std::vector<std::vector<int>> dcls_dims;
        unsigned long long off = 0;
        for (int r = 0; r < R; r++) {
            std::vector<int> d = {sp.dims[3 * r], sp.dims[3 * r + 1], sp.dims[3 * r + 2]};
            for (int a = 0; a < 3; a++) {
                SNH_REQUIRE(d[a] >= 1, SNH_ERR_INVALID, "replica dims must be positive");
                if (lw[a] > d[a]) G.general = 1;
            }
            auto it = dcls_of.find(d);
            if (it == dcls_of.end()) {
                it = dcls_of.emplace(d, (int)dcls_dims.size()).first;
                dcls_dims.push_back(d);
            }
            ReplicaInfo ri{};
            ri.L1 = d[0];
            ri.L2 = d[1];
            ri.L3 = d[2];
            ri.P1 = d[0] + 2 * lw[0];
            ri.P2 = d[1] + 2 * lw[1];
            ri.P3 = d[2] + 2 * lw[2];
            const unsigned long long nc = (unsigned long long)d[0] * d[1] * d[2];
            const unsigned long long pc = (unsigned long long)ri.P1 * ri.P2 * ri.P3;
            ri.ncell = (int)nc;
            ri.pcell = (uint32_t)pc;
            ri.dcls = it->second;
            ri.off = (uint32_t)off;
            ri.inv1 = std::nextafter(1.0f / (float)d[0], 2.0f);
            ri.inv2 = std::nextafter(1.0f / (float)d[1], 2.0f);
            SNH_REQUIRE(nc < (1ull << 24), SNH_ERR_UNSUPPORTED, "a replica must have fewer than 2^24 cells");
            off += pc * na;
            SNH_REQUIRE(off * sizeof(T) < (1ull << 32), SNH_ERR_UNSUPPORTED,
                        "state planes must stay below 4 GiB (32-bit offsets); split the batch");
            reps.push_back(ri);
        }
        total = (uint32_t)off;
        // the tiled sweep reads the three planes of a state through one
        // buffer descriptor with 32-bit byte offsets
        SNH_REQUIRE(!tiled || 3ull * total * sizeof(T) < (1ull << 32), SNH_ERR_UNSUPPORTED,
                    "a tiled batch's state must stay below 4 GiB (32-bit offsets); split the batch");
        // blocks: replica-major, then chunks of kBlock interior cells, then
        // sublattice: consecutive blocks cover the same cells of different
        // sublattices, which is where a site's bonded neighbours live
        std::vector<BlockInfo> blocks;
        rblk.assign(R + 1, 0);
        for (int r = 0; r < R; r++) {
            rblk[r] = (int)blocks.size();
            for (int c0 = 0; c0 < reps[r].ncell; c0 += kBlock)
                for (int i = 0; i < na; i++) blocks.push_back({r, i, c0, 0});
        }
        rblk[R] = (int)blocks.size();
        nblocks = (int)blocks.size();

        // bonds: group by source sublattice (gather lists); pad every list to a
        // multiple of the gather chunk with zero couplings of the site to itself
        std::vector<int> order(nb);
        for (int b = 0; b < nb; b++) order[b] = b;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sp.bond[6 * a] < sp.bond[6 * b]; });
        std::vector<int> count(na, 0);
        for (int b = 0; b < nb; b++) count[sp.bond[6 * b]]++;
        long long slots6 = 0, slots9 = 0, slots18 = 0;
        for (int i = 0; i < na; i++) {
            slots6 += (count[i] + 5) / 6 * 6;
            slots9 += (count[i] + 8) / 9 * 9;
            slots18 += (count[i] + 17) / 18 * 18;
        }
        // neighbours per batch of loads: the chunk with the fewest padded slots
        // (ties to the larger chunk: fewer round trips); SU(N) kernels use 6
        int CH = 6;
        if (N == 0) CH = slots18 <= slots6 ? 18 : (slots9 <= slots6 ? 9 : 6);
        C.chunk = CH;
        std::vector<int> bstart(na + 1, 0);
        for (int i = 0; i < na; i++) bstart[i + 1] = bstart[i] + (count[i] + CH - 1) / CH * CH;
        const int nbp = bstart[na];  // padded bond count
        std::vector<int> bj(std::max(nbp, 1)), bflag(std::max(nbp, 1));
        std::vector<int> bsrc(std::max(nbp, 1), -1);  // original bond index, -1 for padding
        std::vector<T> bJ(9 * (size_t)std::max(nbp, 1), (T)0), bJ0(std::max(nbp, 1), (T)0),
            bQ(25 * (size_t)std::max(nbp, 1), (T)0), bsc(std::max(nbp, 1), (T)0);
        // general (A, B) terms of every padded slot, in slot order
        const bool have_gen = sp.gen_start != nullptr && N > 0;
        SNH_REQUIRE(N > 0 || !sp.gen_start || sp.gen_start[sp.nbonds] == sp.gen_start[0], SNH_ERR_INVALID,
                    "general pair couplings require :SUN mode (in dipole mode the most general coupling is "
                    "bilinear plus biquadratic)");
        SNH_REQUIRE(!have_gen || (sp.gen_A && sp.gen_B), SNH_ERR_INVALID, "general coupling terms missing");
        std::vector<int> gstart(std::max(nbp, 1) + 1, 0);
        std::vector<T> gA, gB;
        {
            std::vector<int> fill(na, 0);
            for (int k = 0; k < nb; k++) {
                const int b = order[k];
                const int* bd = sp.bond + 6 * b;
                const int slot = bstart[bd[0]] + fill[bd[0]]++;
                bsrc[slot] = b;
                bj[slot] = bd[1];
                const double* J = sp.bilin + 9 * b;
                const bool iso = J[1] == 0 && J[2] == 0 && J[3] == 0 && J[5] == 0 && J[6] == 0 && J[7] == 0 &&
                                 J[0] == J[4] && J[0] == J[8];
                const bool bq = sp.has_biquad && sp.has_biquad[b];
                bflag[slot] = (iso ? kBondIso : 0) | (bq ? kBondBiquad : 0) | (bd[5] ? kBondCulled : 0);
                for (int p = 0; p < 9; p++) bJ[9 * (size_t)slot + p] = (T)J[p];
                bJ0[slot] = (T)J[0];
                for (int p = 0; p < 25; p++) bQ[25 * (size_t)slot + p] = bq ? (T)sp.biquad[25 * b + p] : (T)0;
                bsc[slot] = sp.scalar ? (T)sp.scalar[b] : (T)0;
                if (have_gen) {
                    SNH_REQUIRE(sp.gen_start[b] >= 0 && sp.gen_start[b] <= sp.gen_start[b + 1], SNH_ERR_INVALID,
                                "gen_start must be non-decreasing");
                    if (sp.gen_start[b + 1] > sp.gen_start[b]) bflag[slot] |= kBondGeneral;
                }
            }
            if (have_gen)
                for (int slot = 0; slot < nbp; slot++) {
                    gstart[slot] = (int)(gA.size() / (2 * (size_t)N * N));
                    const int b = bsrc[slot];
                    if (b < 0) continue;
                    for (int t = sp.gen_start[b]; t < sp.gen_start[b + 1]; t++)
                        for (size_t q = 0; q < 2 * (size_t)N * N; q++) {
                            gA.push_back((T)sp.gen_A[2 * (size_t)N * N * t + q]);
                            gB.push_back((T)sp.gen_B[2 * (size_t)N * N * t + q]);
                        }
                }
            if (have_gen) gstart[nbp] = (int)(gA.size() / (2 * (size_t)N * N));
            for (int i = 0; i < na; i++)
                for (int slot = bstart[i] + count[i]; slot < bstart[i + 1]; slot++) {
                    bj[slot] = i;
                    bflag[slot] = kBondIso | kBondCulled;
                }
        }
        bsrc_h = bsrc;
        bflag_h = bflag;
        nb_in = sp.nbonds;
        bond_h.assign(sp.bond, sp.bond + 6 * (size_t)sp.nbonds);
        for (int b = 0; b < sp.nbonds; b++) any_biquad |= sp.has_biquad && sp.has_biquad[b];
        // general couplings, like biquadratic ones, need the neighbours' kets
        for (int k = 0; k < nbp; k++) any_biquad |= (bflag[k] & kBondGeneral) != 0;
        nb = nbp;
        C.iso = 1;
        for (int k = 0; k < nbp; k++)
            if (!(bflag[k] & kBondIso) || (bflag[k] & (kBondBiquad | kBondGeneral))) C.iso = 0;
        // padded-index offset of every bond, per dims class
        std::vector<int> bdelta(std::max<size_t>(1, dcls_dims.size() * nbp), 0);
        for (size_t d = 0; d < dcls_dims.size(); d++) {
            const long long P1 = dcls_dims[d][0] + 2 * lw[0], P2 = dcls_dims[d][1] + 2 * lw[1],
                            P3 = dcls_dims[d][2] + 2 * lw[2];
            for (int k = 0; k < nbp; k++) {
                // padding is the site itself; so is every bond of a tiled
                // layout, whose ghost-padded gather must never run
                if (bsrc[k] < 0 || tiled) continue;
                const int* bd = sp.bond + 6 * bsrc[k];
                const long long v = bd[2] + P1 * (bd[3] + P2 * (long long)bd[4]) + P1 * P2 * P3 * (bd[1] - bd[0]);
                bdelta[d * nbp + k] = (int)v;
            }
        }

        // per-sublattice tables
        std::vector<T> kap(na), zee(3 * na), stev(28 * na), gm(9 * na), fld(3 * na), q2(9 * na, (T)0);
        std::vector<int> kmax(na, 0);
        for (int i = 0; i < na; i++) {
            kap[i] = (T)sp.kappa[i];
            for (int p = 0; p < 9; p++) gm[9 * i + p] = (T)sp.g[9 * i + p];
            for (int p = 0; p < 3; p++) fld[3 * i + p] = (T)sp.field[3 * i + p];
            for (int b = 0; b < 3; b++) {
                double z = 0;
                for (int a = 0; a < 3; a++) z += sp.g[9 * i + 3 * a + b] * sp.field[3 * i + a];
                zee[3 * i + b] = (T)z;
            }
            if (N == 0) {
                SNH_REQUIRE(sp.stevens && sp.kmax, SNH_ERR_INVALID, ":dipole mode needs Stevens coefficients");
                int km = sp.kmax[i];
                SNH_REQUIRE(km == 0 || km == 2 || km == 4 || km == 6, SNH_ERR_INVALID, "kmax must be 0, 2, 4 or 6");
                kmax[i] = km;
                for (int p = 0; p < 28; p++) stev[28 * i + p] = (T)sp.stevens[28 * i + p];
                // kmax = 2: gradient 2 Q2 S of c2 . (x^2-y^2, zx, 3z^2-X, zy, 2xy),
                // |S|^2 = X held constant (OnsiteCoupling.jl:316-425)
                const double* c2 = sp.stevens + 28 * i + 1;
                const double Q[9] = {2 * c2[0], 2 * c2[4], c2[1], 2 * c2[4], -2 * c2[0], c2[3], c2[1], c2[3], 6 * c2[2]};
                for (int p = 0; p < 9; p++) q2[9 * i + p] = (T)Q[p];
            }
        }
        kmax_h = kmax;
        // SNH_LEAN=0 keeps the general tiled kernels (for comparisons)
        C.lean_ok = env_int("SNH_LEAN", 1) != 0;
        for (int k : kmax)
            if (k > 2) C.lean_ok = 0;
        std::vector<T> lam, qops;
        if (N > 0) {
            SNH_REQUIRE(sp.Lambda, SNH_ERR_INVALID, ":SUN mode needs the onsite matrices Lambda");
            lam.resize((size_t)na * N * N * 2);
            for (size_t k = 0; k < lam.size(); k++) lam[k] = (T)sp.Lambda[k];
            auto q = quadrupole_ops(N);
            qops.assign(q.begin(), q.end());
        } else {
            lam.assign(2, (T)0);
            qops.assign(2, (T)0);
        }

        C.blocks = own(dupload(blocks));
        C.reps = own(dupload(reps));
        C.bstart = own(dupload(bstart));
        C.bj = own(dupload(bj));
        C.bflag = own(dupload(bflag));
        if (gA.empty()) gA.assign(2, (T)0), gB.assign(2, (T)0);
        C.gstart = own(dupload(gstart));
        C.gA = own(dupload(gA));
        C.gB = own(dupload(gB));
        C.bdelta = own(dupload(bdelta));
        C.bJ0 = own(dupload(bJ0));
        C.bJ = own(dupload(bJ));
        C.bQ = own(dupload(bQ));
        C.bscalar = own(dupload(bsc));
        C.kappa = own(dupload(kap));
        C.zee = own(dupload(zee));
        C.stev = own(dupload(stev));
        C.stev_q2 = own(dupload(q2));
        C.kmax = own(dupload(kmax));
        C.gmat = own(dupload(gm));
        C.field = own(dupload(fld));
        C.lam = own(dupload(lam));
        C.quad_ops = own(dupload(qops));
        C.na = na;
        C.nb = nb;
        C.N = N;
        C.total = total;
        C.G = G;
        d_rblk = own(dupload(rblk));

        X = own(dalloc<T>((size_t)planes * total));
        Xp = own(dalloc<T>((size_t)planes * total));
        Xo = own(dalloc<T>((size_t)planes * total));
        SNH_HIP(hipMemset(X, 0, sizeof(T) * planes * (size_t)total));
        if (tiled) {
            // lean FP32 kernels over isotropic bonds run both stages one site
            // per thread on the model's first tile where it offers that plan
            apply_plan(tile_cands[default_correct_plan()], 0);
            apply_plan(tile_cands[default_correct_plan()], 1);
        }
        if (!partial) {
            partial_cap = std::max<size_t>(nblocks, 1);
            partial = own(dalloc<double>(partial_cap));
        }
        d_out = own(dalloc<double>(R));
        conv_iter = own(dalloc<int>(R));
        atol2 = own(dalloc<double>(R));
        nscale = own(dalloc<T>(R));
        SNH_HIP(hipHostMalloc((void**)&pinned, sizeof(double) * R, hipHostMallocDefault));
        SNH_HIP(hipHostMalloc((void**)&pinned_i, sizeof(int) * R, hipHostMallocDefault));
        if (ewald) build_ewald(sp);
        long long max_sites = 0;
        for (const auto& ri : reps) max_sites = std::max(max_sites, (long long)ri.ncell * na);
        // measured crossover: faster up to one site per thread of the
        // workgroup (64 sites: 5.8 vs 8.8 us per Langevin step), slower beyond
        // (4096 sites: 64 vs 10.7 us)
        persist = N == 0 && !ewald && !tiled && env_int("SNH_PERSIST", 1) != 0 &&
                  max_sites <= env_int("SNH_PERSIST_MAX_SITES", kPersist);
        mid_sync = env_int("SNH_MID_SYNC", 0) != 0;
        const int pair_env = env_int("SNH_PAIR", 1);
        if (tiled && !ewald && R >= 2 && pair_env != 0) {
            // split the replicas into two groups of about equal sites
            long long all = 0, acc = 0;
            for (int r = 0; r < R; r++) all += replica_sites(r);
            int h = 1;
            for (int r = 0; r < R - 1; r++) {
                acc += replica_sites(r);
                h = r + 1;
                if (2 * acc >= all) break;
            }
            pair_split = h;
            pair_on = pair_env == 2;
            SNH_HIP(hipStreamCreate(&st2));
            for (hipEvent_t& e : pev) SNH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        }
    }

    // The part of a tile plan's work list that covers replicas [r0, r1).
    static TileArgs plan_part(const TilePlan& P, int r0, int r1)
    {
        TileArgs A = P.A;
        A.work = P.A.work + P.rwork[r0];
        A.nwork = P.rwork[r1] - P.rwork[r0];
        return A;
    }

    // Steps [s0, s1) of the tiled Langevin path as two replica groups:
    // group 0 on st, group 1 on st2; group 1's predictor waits for group 0's,
    // and group 0's next predictor for group 1's, so each corrector overlaps
    // the other group's predictor.  Per-site arithmetic is that of
    // langevin_step, so results are identical.  The correctors write to Sout
    // (the state X, or a scratch buffer when timing).
    void langevin_pairs(NoiseArgs<T> nz, T tdt, T tl, uint64_t step0, int s0, int s1, T* Sout)
    {
        const TileArgs p0 = plan_part(tp, 0, pair_split), p1 = plan_part(tp, pair_split, R);
        const TileArgs c0 = plan_part(tc, 0, pair_split), c1 = plan_part(tc, pair_split, R);
        SNH_HIP(hipEventRecord(pev[0], st));
        SNH_HIP(hipStreamWaitEvent(st2, pev[0], 0));
        for (int s = s0; s < s1; s++) {
            nz.step = step0 + (uint64_t)s;
            if (s > s0) SNH_HIP(hipStreamWaitEvent(st, pev[2], 0));
            launch_tile_predict<T>(C, p0, tp.threads, X, Xp, Xo, nullptr, nz, tdt, tl, st);
            SNH_HIP(hipEventRecord(pev[1], st));
            launch_tile_correct<T>(C, c0, tc.threads, Xp, Xo, Sout, nullptr, tdt, tl, st);
            SNH_HIP(hipStreamWaitEvent(st2, pev[1], 0));
            launch_tile_predict<T>(C, p1, tp.threads, X, Xp, Xo, nullptr, nz, tdt, tl, st2);
            SNH_HIP(hipEventRecord(pev[2], st2));
            launch_tile_correct<T>(C, c1, tc.threads, Xp, Xo, Sout, nullptr, tdt, tl, st2);
        }
        SNH_HIP(hipEventRecord(pev[0], st2));
        SNH_HIP(hipStreamWaitEvent(st, pev[0], 0));
    }

    // Parameters of the real-space tensor of an L1 x L2 x L3 supercell
    // (Ewald.jl:61-91).
    static EwaldAParams ewald_params(const snh_spec& sp, int na, const int L[3])
    {
        EwaldAParams P{};
        for (int p = 0; p < 9; p++) P.lv[p] = sp.latvecs[p];
        for (int rr = 0; rr < 3; rr++)
            for (int c = 0; c < 3; c++) P.SL[3 * rr + c] = sp.latvecs[3 * rr + c] * L[c];
        const double* S = P.SL;
        double det = S[0] * (S[4] * S[8] - S[5] * S[7]) - S[1] * (S[3] * S[8] - S[5] * S[6]) +
                     S[2] * (S[3] * S[7] - S[4] * S[6]);
        SNH_REQUIRE(std::fabs(det) > 0, SNH_ERR_INVALID, "singular lattice vectors");
        double inv[9] = {(S[4] * S[8] - S[5] * S[7]) / det,  -(S[1] * S[8] - S[2] * S[7]) / det,
                         (S[1] * S[5] - S[2] * S[4]) / det,  -(S[3] * S[8] - S[5] * S[6]) / det,
                         (S[0] * S[8] - S[2] * S[6]) / det,  -(S[0] * S[5] - S[2] * S[3]) / det,
                         (S[3] * S[7] - S[4] * S[6]) / det,  -(S[0] * S[7] - S[1] * S[6]) / det,
                         (S[0] * S[4] - S[1] * S[3]) / det};
        for (int rr = 0; rr < 3; rr++)
            for (int c = 0; c < 3; c++) P.RL[3 * rr + c] = 2 * M_PI * inv[3 * c + rr];
        for (int p = 0; p < 9; p++) P.demag[p] = sp.demag[p];
        P.V = det;
        P.sigma = std::cbrt(det) / 3;
        P.rmax = 6 * std::sqrt(2.0) * P.sigma;
        P.kmax = 6 * std::sqrt(2.0) / P.sigma;
        for (int ax = 0; ax < 3; ax++) {
            double av[3] = {P.SL[ax], P.SL[3 + ax], P.SL[6 + ax]};
            double bv[3] = {P.RL[ax], P.RL[3 + ax], P.RL[6 + ax]};
            double an = std::sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
            double bn = std::sqrt(bv[0] * bv[0] + bv[1] * bv[1] + bv[2] * bv[2]);
            double adb = av[0] * bv[0] + av[1] * bv[1] + av[2] * bv[2];
            P.nmax[ax] = (int)std::nearbyint(P.rmax / (adb / bn) + 1e-6) + 1;  // round(): ties to even
            P.mmax[ax] = (int)std::nearbyint(P.kmax / (adb / an) + 1e-6);
        }
        P.mu0_muB2 = sp.mu0_muB2;
        P.L1 = L[0];
        P.L2 = L[1];
        P.L3 = L[2];
        P.na = na;
        return P;
    }

    void build_ewald(const snh_spec& sp)
    {
        SNH_REQUIRE(sp.positions, SNH_ERR_INVALID, "Ewald needs atom positions");
        ew_pos.assign(sp.positions, sp.positions + 3 * na);
        int dev_ = 0;
        SNH_HIP(hipGetDevice(&dev_));
        hipDeviceProp_t prop;
        SNH_HIP(hipGetDeviceProperties(&prop, dev_));
        const size_t lim = prop.sharedMemPerBlock;
        const char* mode = std::getenv("SNH_EWALD_FFT");
        const bool force_lib = mode && std::string(mode) == "lib";
        double* d_pos = dupload(ew_pos);
        own(d_pos);
        for (int r = 0; r < R;) {
            int r1 = r + 1;
            while (r1 < R && reps[r1].L1 == reps[r].L1 && reps[r1].L2 == reps[r].L2 && reps[r1].L3 == reps[r].L3) r1++;
            EwaldGroup g;
            g.r0 = r;
            g.nr = r1 - r;
            const int L[3] = {reps[r].L1, reps[r].L2, reps[r].L3};
            EwaldGeom& EG = g.EG;
            EG.L1 = L[0];
            EG.L2 = L[1];
            EG.L3 = L[2];
            EG.H1 = L[0] / 2 + 1;
            EG.na = na;
            EG.ncell = (long long)L[0] * L[1] * L[2];
            EG.nk = (long long)EG.H1 * L[1] * L[2];
            EG.P1 = reps[r].P1;
            EG.P2 = reps[r].P2;
            EG.pcell = reps[r].pcell;
            EG.w1 = G.w1;
            EG.w2 = G.w2;
            EG.w3 = G.w3;
            // one workgroup per field transforms grids whose FP64 transform of
            // the tensor A (the largest) fits its LDS; larger grids, or any
            // under SNH_EWALD_FFT=lib, go through hipFFT
            const bool fits = L[0] <= kLmax && L[1] <= kLmax && L[2] <= kLmax &&
                              ewald_fft_lds_bytes(EG, sizeof(double)) <= lim;
            EG.plain = !fits || force_lib ? 1 : 0;
            // few replicas: transforms split by axis (SNH_EWALD_SLAB=0 never,
            // 2 up to kSlabMaxReplicas replicas)
            const int slab_mode = env_int("SNH_EWALD_SLAB", 1);
            EG.slab = !EG.plain && slab_mode != 0 &&
                      g.nr <= (slab_mode == 2 ? kSlabMaxReplicas : kEwaldSlabReplicas) &&
                      ewald_slab_plane_lds_bytes(EG, sizeof(T)) <= lim &&
                      ewald_slab_column_lds_bytes(EG, g.nr, sizeof(T)) <= lim ? 1 : 0;
            const size_t lc = ewald_use_mfma(EG, g.nr, sizeof(T)) ? ewald_mfma_lds_bytes(EG)
                                                                  : (size_t)g.nr * 3 * na * sizeof(T) * 2;
            SNH_REQUIRE(lc <= lim, SNH_ERR_UNSUPPORTED, "too many replicas x sublattices for the Ewald contraction");
            g.P = ewald_params(sp, na, L);
            double* d_A = dalloc<double>((size_t)na * na * EG.ncell * 9);
            const int n3 = 3 * na;
            g.M = own(dalloc<char>((size_t)EG.nk * n3 * n3 * 2 * sizeof(T)));
            launch_ewald_A(g.P, d_pos, d_A, st);
            std::string err;
            if (EG.plain)
                err = launch_ewald_fwdA_plain<T>(EG, d_A, g.M, st);
            else
                launch_ewald_fwdA<T>(EG, d_A, g.M, st);
            SNH_HIP(hipGetLastError());
            SNH_HIP(hipStreamSynchronize(st));
            SNH_HIP(hipFree(d_A));
            SNH_REQUIRE(err.empty(), SNH_ERR_HIP, "Ewald tensor: " + err);
            g.Fmu = own(dalloc<char>(ewald_spectrum_bytes(EG, g.nr, sizeof(T))));
            g.Fphi = own(dalloc<char>(ewald_spectrum_bytes(EG, g.nr, sizeof(T))));
            if (EG.plain) {
                g.lib = ewald_lib_create(EG, g.nr, sizeof(T), err);
                SNH_REQUIRE(g.lib, SNH_ERR_HIP, "Ewald FFT: " + err);
            }
            ewg.push_back(g);
            r = r1;
        }
        phi1 = own(dalloc<T>(3 * (size_t)total));
        phi2 = own(dalloc<T>(3 * (size_t)total));
        if (N > 0) Sd = own(dalloc<T>(3 * (size_t)total));
    }

    bool ewald_lib_any() const
    {
        for (const auto& g : ewg)
            if (g.EG.plain) return true;
        return false;
    }

    // --- timing helpers
    hipEvent_t tbegin()
    {
        if (!timer.enabled) return nullptr;
        hipEvent_t e = timer.get();
        SNH_HIP(hipEventRecord(e, st));
        return e;
    }
    void tend(const char* fam, hipEvent_t e0)
    {
        if (!e0) return;
        hipEvent_t e1 = timer.get();
        SNH_HIP(hipEventRecord(e1, st));
        timer.events.push_back({fam, {e0, e1}});
    }

    // Ewald potential phi of state planes Xs (dipoles; expected dipoles for
    // SU(N), computed here unless the caller holds them in dip).
    void ewald_field(const T* Xs, T* phi, const T* dip = nullptr, bool timed = true)
    {
        const T* S = dip ? dip : Xs;
        if (N > 0 && !dip) {
            launch_sun_expected_spin<T>(C, nblocks, Xs, Sd, st);
            S = Sd;
        }
        hipEvent_t e = timed ? tbegin() : nullptr;
        for (auto& g : ewg)
            SNH_REQUIRE(launch_ewald_field<T>(g.EG, g.nr, C.reps + g.r0, total, C.gmat, S, g.M, g.Fmu, g.Fphi, phi,
                                              g.lib, st),
                        SNH_ERR_HIP, "hipFFT execution failed");
        tend("ewald_field", e);
    }

    // Calls f(padded site, interior site) for every padded site of every
    // replica, interior site = the periodic source of the (ghost) position.
    template <typename F>
    void for_each_padded(F f) const
    {
        long long isite0 = 0;
        for (int r = 0; r < R; r++) {
            const ReplicaInfo& ri = reps[r];
            const int wv[3] = {G.w1, G.w2, G.w3};
            for (int i = 0; i < na; i++)
                for (int p3 = 0; p3 < ri.P3; p3++)
                    for (int p2 = 0; p2 < ri.P2; p2++)
                        for (int p1 = 0; p1 < ri.P1; p1++) {
                            const int c1 = ((p1 - wv[0]) % ri.L1 + ri.L1) % ri.L1;
                            const int c2 = ((p2 - wv[1]) % ri.L2 + ri.L2) % ri.L2;
                            const int c3 = ((p3 - wv[2]) % ri.L3 + ri.L3) % ri.L3;
                            const uint32_t g = ri.off + (uint32_t)(p1 + ri.P1 * (p2 + ri.P2 * p3)) + ri.pcell * i;
                            const long long s = isite0 + c1 + (long long)ri.L1 * (c2 + (long long)ri.L2 * c3) +
                                                (long long)ri.ncell * i;
                            const bool interior = p1 - wv[0] == c1 && p2 - wv[1] == c2 && p3 - wv[2] == c3;
                            f(g, s, interior);
                        }
            isite0 += (long long)ri.ncell * na;
        }
    }

    long long interior_sites() const
    {
        long long n = 0;
        for (const auto& ri : reps) n += (long long)ri.ncell * na;
        return n;
    }

    void download_planes(const T* P, int np, double* x)
    {
        std::vector<T> h((size_t)np * total);
        SNH_HIP(hipMemcpyAsync(h.data(), P, sizeof(T) * h.size(), hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        for_each_padded([&](uint32_t g, long long s, bool interior) {
            if (!interior) return;
            for (int p = 0; p < np; p++) x[(size_t)np * s + p] = (double)h[(size_t)p * total + g];
        });
    }

    int64_t total_sites() const override { return interior_sites(); }
    int replicas() const override { return R; }
    int64_t replica_sites(int r) const override { return (int64_t)reps[r].ncell * na; }

    void set_state(const double* x, bool coherent) override
    {
        dz_valid = false;
        SNH_REQUIRE(coherent == (N > 0), SNH_ERR_INVALID,
                    coherent ? "coherent states require :SUN mode" : "in :SUN mode set coherent states");
        std::vector<T> h((size_t)planes * total, (T)0);
        for_each_padded([&](uint32_t g, long long s, bool) {
            for (int p = 0; p < planes; p++) h[(size_t)p * total + g] = (T)x[(size_t)planes * s + p];
        });
        SNH_HIP(hipMemcpyAsync(X, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, st));
        SNH_HIP(hipStreamSynchronize(st));
    }

    void get_state(double* x, bool coherent) override
    {
        if (coherent || N == 0) {
            SNH_REQUIRE(!coherent || N > 0, SNH_ERR_INVALID, "coherent states require :SUN mode");
            download_planes(X, planes, x);
        } else {
            // dipoles of a coherent state: expected_spin(Z) (Integrators.jl:354)
            launch_sun_expected_spin<T>(C, nblocks, X, Xp, st);  // stage buffer is free between steps
            SNH_HIP(hipGetLastError());
            download_planes(Xp, 3, x);
        }
    }

    // The midpoint's buffers, each allocated on first need: S'' ping-pong B1/B2,
    // and Shat (:dipole, ping-pong Bh/Bh2) or Zbar (SU(N) + Ewald) in Bh.
    void ensure_midpoint_buffers()
    {
        const size_t n = (size_t)planes * total;
        if (!B1) B1 = own(dalloc<T>(n));
        if (!B2) B2 = own(dalloc<T>(n));
        if ((ewald || N == 0) && !Bh) Bh = own(dalloc<T>(n));
        if (N == 0 && !Bh2) Bh2 = own(dalloc<T>(n));
        if (!last_it) last_it = own(dalloc<int>(R));
        if (tiled && N == 0 && !Bmu) Bmu = own(dalloc<T>(2 * (size_t)total));
        SNH_REQUIRE(B1 && B2 && (Bh || (!ewald && N != 0)) && (Bh2 || N != 0) && last_it && (Bmu || !tiled || N != 0),
                    SNH_ERR_HIP,
                    "midpoint buffers missing");
    }

    // Makes P the tiled plan of the predictor side (which = 0) or of the
    // corrector side (which = 1): device work list, LDS bond offsets per
    // ring phase, per-replica workgroup ranges, and room for its partial sums.
    // The plan of both stages (and of the midpoint iteration) before any
    // autotuning: the cost model's first, except that lean FP32 kernels over
    // isotropic bonds take the same tile with one site per thread where the
    // model offers it (their one-site forms fit two 1024-thread workgroups
    // per CU: the corrector in 64 registers, the predictor in 60).
    size_t default_correct_plan() const
    {
        if (sizeof(T) != 4 || !C.lean_ok || !C.iso || ewald || env_int("SNH_TILE_SP", 0)) return 0;
        const TilePlan& b = tile_cands[0];
        for (size_t k = 1; k < tile_cands.size(); k++) {
            const TilePlan& a = tile_cands[k];
            if (a.A.sp == 1 && b.A.sp == 2 && a.A.T1 == b.A.T1 && a.A.T2 == b.A.T2 && a.work.size() == b.work.size())
                return k;
        }
        return 0;
    }

    void apply_plan(const TilePlan& P, int which)
    {
        if (st) SNH_HIP(hipStreamSynchronize(st));
        for (void* p : plan_mem[which]) (void)hipFree(p);
        plan_mem[which].clear();
        TilePlan& D = which ? tc : tp;
        D = P;
        const std::vector<int> ldsoff = tile_lds_offsets(D.A, na, bond_h.data(), bsrc_h, sizeof(T));
        TileWork* w = dupload(D.work);
        int* o = dupload(ldsoff);
        int* rb = dupload(D.rwork);
        plan_mem[which] = {w, o, rb};
        D.A.work = w;
        D.A.ldsoff = o;
        (which ? d_tcblk : d_trblk) = rb;
        const size_t need = std::max<size_t>(nblocks, D.work.size());
        if (need > partial_cap) {
            if (partial) {
                owned.erase(std::remove(owned.begin(), owned.end(), (void*)partial), owned.end());
                (void)hipFree(partial);
            }
            partial = own(dalloc<double>(need));
            partial_cap = need;
        }
        drop_graph();  // captured launches carry the old plan
    }

    // Times up to `candidates` of the cost model's plans on this handle, the
    // Heun predictor and corrector separately (one warm-up, then `launches`
    // launches each; the corrector writes into a scratch buffer so the state
    // is untouched), and keeps the fastest for each.  A later-ranked plan
    // must win by 2% to displace an earlier one (timing noise).
    void autotune(int candidates, int launches, double* best_ms) override
    {
        SNH_REQUIRE(candidates >= 1 && launches >= 1, SNH_ERR_INVALID, "candidates and launches must be positive");
        if (!tiled) {
            if (best_ms) *best_ms = 0;
            return;
        }
        ensure_midpoint_buffers();  // B1 is the corrector's scratch output here
        NoiseArgs<T> nz{};
        nz.mode = 2;  // device noise: the predictor's full work
        nz.scale = nscale;
        const T dt = (T)0.01, lam = (T)0.1;
        hipEvent_t e0, e1;
        SNH_HIP(hipEventCreate(&e0));
        SNH_HIP(hipEventCreate(&e1));
        auto timed = [&](auto&& launch) {
            launch();
            SNH_HIP(hipEventRecord(e0, st));
            for (int k = 0; k < launches; k++) launch();
            SNH_HIP(hipEventRecord(e1, st));
            SNH_HIP(hipEventSynchronize(e1));
            float ms = 0;
            SNH_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms / launches;
        };
        const int nc = std::min<int>(candidates, (int)tile_cands.size());
        int best[2] = {0, 0};
        float best_t[2] = {0, 0};
        for (int c = 0; c < nc; c++) {
            apply_plan(tile_cands[c], 0);
            apply_plan(tile_cands[c], 1);
            const float t[2] = {
                timed([&] { launch_tile_predict<T>(C, tp.A, tp.threads, X, Xp, Xo, nullptr, nz, dt, lam, st); }),
                timed([&] { launch_tile_correct<T>(C, tc.A, tc.threads, Xp, Xo, B1, nullptr, dt, lam, st); })};
            for (int k = 0; k < 2; k++)
                if (c == 0 || t[k] < 0.98f * best_t[k]) {
                    best_t[k] = t[k];
                    best[k] = c;
                }
            if (env_int("SNH_AUTOTUNE_VERBOSE", 0))
                std::fprintf(stderr, "[snh_autotune] plan %d: T1 %d T2 %d sp %d threads %d workgroups %d lds %zu: "
                                     "predictor %.4f ms, corrector %.4f ms\n",
                             c, tp.A.T1, tp.A.T2, tp.A.sp, tp.threads, tp.A.nwork, tp.lds, t[0], t[1]);
        }
        apply_plan(tile_cands[best[0]], 0);
        apply_plan(tile_cands[best[1]], 1);
        tuned_index[0] = best[0];
        tuned_index[1] = best[1];
        if (best_ms) *best_ms = best_t[0] + best_t[1];
        // two replica groups in anti-phase pay off for some plan pairs and
        // not for others: keep them when whole steps (correctors into the
        // scratch buffer) are 2% faster than on one stream
        if (pair_split > 0 && env_int("SNH_PAIR", 1) == 1) {
            const int nstep = std::max(launches, 8);
            auto steps = [&](auto&& run) {  // a warm-up run, then one timed run of nstep steps
                run();
                SNH_HIP(hipEventRecord(e0, st));
                run();
                SNH_HIP(hipEventRecord(e1, st));
                SNH_HIP(hipEventSynchronize(e1));
                float ms = 0;
                SNH_HIP(hipEventElapsedTime(&ms, e0, e1));
                return ms;
            };
            const float one = steps([&] {
                for (int k = 0; k < nstep; k++) {
                    launch_tile_predict<T>(C, tp.A, tp.threads, X, Xp, Xo, nullptr, nz, dt, lam, st);
                    launch_tile_correct<T>(C, tc.A, tc.threads, Xp, Xo, B1, nullptr, dt, lam, st);
                }
            });
            const float two = steps([&] { langevin_pairs(nz, dt, lam, 0, 0, nstep, B1); });
            pair_on = two < 0.98f * one;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    }

    // SampledCorrelations(sys; measure, energies, dt) buffers
    // (SampledCorrelations.jl:194-268): T = n_all_omega time slots,
    // observables o_a,i (A = o . S at atom i), correlation pairs (alpha, beta).
    void corr_init(int nobs, const double* obs, int ncorr, const int32_t* pairs, int nt, int measperiod,
                   int window) override
    {
        SNH_REQUIRE(R == 1, SNH_ERR_UNSUPPORTED, "sampled correlations support one replica");
        SNH_REQUIRE(nobs >= 1 && obs && ncorr >= 1 && pairs, SNH_ERR_INVALID, "observables and pairs required");
        SNH_REQUIRE(nt >= 2 && nt % 2 == 0 && nt <= 4096, SNH_ERR_INVALID, "n_all_omega must be even, 2..4096");
        SNH_REQUIRE(measperiod >= 1, SNH_ERR_INVALID, "measperiod must be positive");
        const ReplicaInfo& ri = reps[0];
        const int L[3] = {ri.L1, ri.L2, ri.L3};
        for (int c = 0; c < 2 * ncorr; c++)
            SNH_REQUIRE(pairs[c] >= 0 && pairs[c] < nobs, SNH_ERR_INVALID, "correlation pair out of range");
        if (st) SNH_HIP(hipStreamSynchronize(st));
        for (void* p : corr.mem) (void)hipFree(p);
        corr_lib_destroy(corr.lib);
        corr = Corr{};
        // axes up to 32 cells: line transforms in registers; longer axes, or
        // any under SNH_CORR_FFT=lib: a batched hipFFT plan
        const char* fmode = std::getenv("SNH_CORR_FFT");
        if (L[0] > kLmax || L[1] > kLmax || L[2] > kLmax || (fmode && std::string(fmode) == "lib")) {
            std::string err;
            corr.lib = corr_lib_create((long long)(nt / 2 + 1) * nobs * na, L[0], L[1], L[2], err);
            SNH_REQUIRE(corr.lib, SNH_ERR_HIP, "sampled correlations: " + err);
        }
        corr.nobs = nobs;
        corr.ncorr = ncorr;
        corr.nt = nt;
        corr.measperiod = measperiod;
        corr.ncell = ri.ncell;
        for (int a = 0; a < 3; a++) corr.L[a] = L[a];
        corr.pairs.assign(pairs, pairs + 2 * ncorr);
        const size_t cplx = 2 * sizeof(double);
        const size_t nsamp = (size_t)nt * nobs * na * corr.ncell, nchan = (size_t)nt * na * corr.ncell;
        auto alloc = [&](size_t bytes) {
            void* p = dalloc<char>(bytes);
            corr.mem.push_back(p);
            return p;
        };
        corr.obs = static_cast<double*>(alloc(sizeof(double) * 3 * nobs * na));
        SNH_HIP(hipMemcpy(corr.obs, obs, sizeof(double) * 3 * nobs * na, hipMemcpyHostToDevice));
        corr.samp = alloc(nsamp * cplx);
        corr.samp2 = alloc(nsamp * cplx);
        corr.P = alloc(nchan * cplx);
        corr.Y = alloc(nchan * cplx);
        corr.data = alloc((size_t)ncorr * na * nchan * cplx);
        SNH_HIP(hipMemset(corr.data, 0, (size_t)ncorr * na * nchan * cplx));
        // Time offsets fftfreq(T, T) = 0, 1, ..., T/2 - 1, -T/2, ..., -1 of a
        // correlation of two signals of Tsnap = T/2 + 1 snapshots: n_contrib =
        // Tsnap - |offset| estimates each; optional cos^2 window over the slots
        // (CorrelationSampling.jl:51-117); 1/T of the normalised inverse FFT.
        std::vector<double> post(nt);
        const int Tsnap = nt / 2 + 1;
        for (int t = 0; t < nt; t++) {
            const int off = t < nt / 2 ? t : t - nt;
            const double c = std::cos(M_PI * t / nt);
            const double w = window ? c * c : 1.0;
            post[t] = w / ((double)nt * (Tsnap - std::abs(off)));
        }
        corr.post = static_cast<double*>(alloc(sizeof(double) * nt));
        SNH_HIP(hipMemcpy(corr.post, post.data(), sizeof(double) * nt, hipMemcpyHostToDevice));
    }

    // add_sample!(sc, sys) = new_sample! + accum_sample!
    // (CorrelationSampling.jl:21-138, :154-175): a trajectory of
    // ImplicitMidpoint(dt) with damping 0 records Tsnap snapshots every
    // measperiod steps (the handle ends at its last state), then the sample's
    // correlations enter the running mean.
    void corr_add_sample(double dt, double tol, int max_iters) override
    {
        SNH_REQUIRE(corr.nt > 0, SNH_ERR_INVALID, "sampled correlations are not initialised");
        const int nt = corr.nt, Tsnap = nt / 2 + 1, nobs = corr.nobs;
        const long long ncell = corr.ncell;
        const size_t nsamp = (size_t)nt * nobs * na * ncell;
        SNH_HIP(hipMemsetAsync(corr.samp, 0, nsamp * 2 * sizeof(double), st));
        const double scale = 1.0 / std::sqrt((double)ncell);  // the 1/sqrt(prod(dims)) of space_fft!
        // observables act on the dipoles: :SUN states give their expected spins
        if (N > 0 && !Sd) Sd = own(dalloc<T>(3 * (size_t)total));
        auto observe = [&](int n) {
            if (N > 0) launch_sun_expected_spin<T>(C, nblocks, X, Sd, st);
            launch_corr_observe<T>(C, nblocks, N > 0 ? Sd : X, nobs, corr.obs, corr.samp, n, scale, st);
        };
        observe(0);
        for (int n = 1; n < Tsnap; n++) {
            step_midpoint(dt, 0.0, tol, max_iters, corr.measperiod, nullptr);
            observe(n);
        }
        // lattice transform of the recorded slots (the padding stays zero)
        if (corr.lib)
            SNH_REQUIRE(corr_lib_exec(corr.lib, corr.samp, st), SNH_ERR_HIP, "hipFFT execution failed");
        else
            launch_corr_space_fft(corr.samp, (long long)Tsnap * nobs * na, corr.L[0], corr.L[1], corr.L[2], st);
        launch_corr_tdft(corr.samp, corr.samp2, (long long)nobs * na * ncell, nt, -1, nullptr, nullptr, 0, st);
        const long long count = ++corr.nsamples;
        const long long nl = (long long)na * ncell;
        const size_t chan = (size_t)nt * nl * 2 * sizeof(double);
        for (int c = 0; c < corr.ncorr; c++)
            for (int i = 0; i < na; i++) {
                launch_corr_product(corr.samp2, corr.P, nt, nobs, na, ncell, corr.pairs[2 * c], i,
                                    corr.pairs[2 * c + 1], st);
                launch_corr_tdft(corr.P, corr.Y, nl, nt, +1, nullptr, corr.post, 0, st);
                launch_corr_tdft(corr.Y, static_cast<char*>(corr.data) + ((size_t)c * na + i) * chan, nl, nt, -1,
                                 nullptr, nullptr, count, st);
            }
        SNH_HIP(hipGetLastError());
    }

    // sc.data in Julia's column-major order data[c, i, j, q1, q2, q3, w]
    // (SampledCorrelations.jl:3), complex as (re, im) pairs.
    void corr_data(double* out, int64_t* nsamples) override
    {
        SNH_REQUIRE(corr.nt > 0, SNH_ERR_INVALID, "sampled correlations are not initialised");
        const int nt = corr.nt, nc = corr.ncorr;
        const long long ncell = corr.ncell;
        const size_t n = (size_t)nc * na * nt * na * ncell;
        std::vector<double> h(2 * n);
        SNH_HIP(hipMemcpyAsync(h.data(), corr.data, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        if (out)
            for (int c = 0; c < nc; c++)
                for (int i = 0; i < na; i++)
                    for (int k = 0; k < nt; k++)
                        for (int j = 0; j < na; j++)
                            for (long long q = 0; q < ncell; q++) {
                                const size_t src = ((((size_t)c * na + i) * nt + k) * na + j) * ncell + q;
                                const size_t dst = c + (size_t)nc * (i + (size_t)na * (j + (size_t)na * (q + ncell * k)));
                                out[2 * dst] = h[2 * src];
                                out[2 * dst + 1] = h[2 * src + 1];
                            }
        if (nsamples) *nsamples = corr.nsamples;
    }

    // Tables of the Metropolis sweeps: colour classes, and the cell offset
    // and target sublattice of every gather slot.
    void mc_prepare()
    {
        if (mc.ready) return;
        std::vector<std::array<int, 3>> dims;
        for (const auto& ri : reps) dims.push_back({ri.L1, ri.L2, ri.L3});
        // a bond that maps a site onto itself (local_energy_change raises)
        for (const auto& d : dims)
            for (int b = 0; b < nb_in; b++) {
                const int32_t* bd = bond_h.data() + 6 * b;
                SNH_REQUIRE(!(bd[0] == bd[1] && bd[2] % d[0] == 0 && bd[3] % d[1] == 0 && bd[4] % d[2] == 0),
                            SNH_ERR_INVALID, "Energy delta for self-interaction not supported");
            }
        mc.col = color_classes(na, nb_in, bond_h.data(), dims);
        // with the Ewald term trials run one at a time: no colouring needed
        SNH_REQUIRE(ewald || mc.col.ncolors > 0, SNH_ERR_UNSUPPORTED,
                    "no colouring of the sites has a period that divides every replica's dims");
        // gather slots: cell offset and target; padding is the site itself (zero coupling)
        std::vector<int> bn(4 * std::max<size_t>(bsrc_h.size(), 1), 0);
        {
            std::vector<int> bst(na + 1, 0);  // slot ranges per sublattice, as uploaded in C.bstart
            SNH_HIP(hipMemcpy(bst.data(), C.bstart, sizeof(int) * (na + 1), hipMemcpyDeviceToHost));
            for (int i = 0; i < na; i++)
                for (int k = bst[i]; k < bst[i + 1]; k++) {
                    const int b = bsrc_h[k];
                    const int32_t* bd = bond_h.data() + 6 * (b < 0 ? 0 : b);
                    const int v[4] = {b < 0 ? 0 : bd[2], b < 0 ? 0 : bd[3], b < 0 ? 0 : bd[4], b < 0 ? i : bd[1]};
                    for (int a = 0; a < 4; a++) bn[4 * (size_t)k + a] = v[a];
                }
        }
        mc.wide = false;
        for (const auto& d : dims)
            for (int b = 0; b < nb_in; b++)
                for (int a = 0; a < 3; a++) mc.wide |= std::abs(bond_h[6 * b + 2 + a]) > d[a];
        // work lists: per colour, per replica, kBlock sites of one class at a time
        const int M = mc.col.m[0] * mc.col.m[1] * mc.col.m[2];
        const int nc = mc.col.ncolors;
        std::vector<McWork> work;
        std::vector<int> start((size_t)nc * (R + 1));
        mc.coff.assign(nc + 1, 0);
        for (int c = 0; c < nc; c++) {
            mc.coff[c] = (int)work.size();
            for (int r = 0; r < R; r++) {
                start[(size_t)c * (R + 1) + r] = (int)work.size();
                const int nq = reps[r].ncell / M;
                for (int cls = 0; cls < na * M; cls++) {
                    if (mc.col.color[cls] != c) continue;
                    for (int q0 = 0; q0 < nq; q0 += kBlock) work.push_back({r, cls, q0, 0});
                }
            }
            start[(size_t)c * (R + 1) + R] = (int)work.size();
        }
        mc.coff[nc] = (int)work.size();
        mc.bn = own(dupload(bn));
        mc.work = own(dupload(work));
        mc.start = own(dupload(start));
        mc.kT = own(dalloc<double>(R));
        mc.acc = own(dalloc<double>(4 * std::max<size_t>(work.size(), (size_t)R)));
        mc.out = own(dalloc<double>(4 * (size_t)R));
        if (ewald) {
            // the real-space tensor (freed after the Fourier transform at
            // creation) is recomputed and kept as the upper triangles of its
            // symmetric 3x3 blocks: every accepted move updates the
            // potential with one of its slices
            for (const auto& g : ewg) {
                const double* D = g.P.demag;
                SNH_REQUIRE(D[1] == D[3] && D[2] == D[6] && D[5] == D[7], SNH_ERR_UNSUPPORTED,
                            "LocalSampler with the Ewald term needs a symmetric demagnetization tensor");
            }
            double* d_pos = dupload(ew_pos);
            own(d_pos);
            for (auto& g : ewg) {
                const size_t nblk = (size_t)na * na * (size_t)g.EG.ncell;
                double* A9 = dalloc<double>(nblk * 9);
                try {
                    launch_ewald_A(g.P, d_pos, A9, st);
                    SNH_HIP(hipGetLastError());
                    SNH_HIP(hipStreamSynchronize(st));
                    std::vector<double> a0(9 * (size_t)na);
                    for (int i = 0; i < na; i++)
                        SNH_HIP(hipMemcpy(a0.data() + 9 * (size_t)i, A9 + ((size_t)(i * na + i) * g.EG.ncell) * 9,
                                          9 * sizeof(double), hipMemcpyDeviceToHost));
                    g.A0 = own(dupload(a0));
                    g.Aew = own(dalloc<double>(nblk * 6));
                    launch_ewald_A_sym(A9, g.Aew, (long long)nblk, st);
                    SNH_HIP(hipGetLastError());
                    SNH_HIP(hipStreamSynchronize(st));
                } catch (...) {
                    (void)hipFree(A9);
                    throw;
                }
                SNH_HIP(hipFree(A9));
            }
            for (int b = 0; b < 2; b++) mc.phi[b] = own(dalloc<double>(3 * (size_t)total));
            mc.moves = own(dalloc<McTrial>((size_t)kMcFftPeriod * R * kMcChunk));
            mc.nmoves = own(dalloc<int>((size_t)kMcFftPeriod * R));
            mc.pprev = own(dalloc<double>((size_t)R * kMcChunk * 3));
        }
        mc.ready = true;
    }

    // step!(sys, LocalSampler(; kT, nsweeps, propose)) for every replica:
    // ceil(2 nsweeps) passes over the colour classes, each trying every site
    // with probability 1/2 (the last pass: the remainder), numbered from
    // 2 sweep0.
    void step_local(const double* kT, int nprop, const int32_t* kinds, const double* mags, const double* weights,
                    uint64_t seed, uint64_t sweep0, double nsweeps, double* dE_dS) override
    {
        SNH_REQUIRE(kT, SNH_ERR_INVALID, "null temperatures");
        SNH_REQUIRE(nprop >= 1 && nprop <= 4 && kinds && weights, SNH_ERR_INVALID, "give 1 to 4 proposals");
        SNH_REQUIRE(std::isfinite(nsweeps) && nsweeps >= 0 && nsweeps < 1e9, SNH_ERR_INVALID,
                    "nsweeps must be in [0, 1e9)");
        for (int r = 0; r < R; r++)
            SNH_REQUIRE(std::isfinite(kT[r]) && kT[r] >= 0, SNH_ERR_INVALID, "kT must be finite and >= 0");
        McArgs A{};
        double wsum = 0;
        for (int p = 0; p < nprop; p++) {
            SNH_REQUIRE(kinds[p] >= 0 && kinds[p] <= 2, SNH_ERR_INVALID, "proposal kind must be 0, 1 or 2");
            const double mg = mags ? mags[p] : 0.0;
            SNH_REQUIRE(kinds[p] != 2 || mg >= 0, SNH_ERR_INVALID, "perturbation magnitude must be >= 0");
            SNH_REQUIRE(std::isfinite(weights[p]) && weights[p] >= 0, SNH_ERR_INVALID,
                        "proposal weights must be finite and >= 0");
            A.kind[p] = kinds[p];
            A.mag[p] = mg;
            wsum += weights[p];
        }
        SNH_REQUIRE(wsum > 0, SNH_ERR_INVALID, "proposal weights must not all be zero");
        double c = 0;
        for (int p = 0; p < nprop; p++) {
            c += weights[p] / wsum;
            A.cum[p] = c;
        }
        mc_prepare();
        // the sweeps rewrite the state: the expected-dipole planes carried
        // between SU(N) Langevin calls no longer describe it
        dz_valid = false;
        A.nprop = nprop;
        A.bn = mc.bn;
        A.kT = mc.kT;
        A.wide = mc.wide ? 1 : 0;
        A.seed = seed;
        A.m1 = mc.col.m[0];
        A.m2 = mc.col.m[1];
        A.m3 = mc.col.m[2];
        SNH_HIP(hipMemcpyAsync(mc.kT, kT, sizeof(double) * R, hipMemcpyHostToDevice, st));
        if (ewald) {
            // ceil(nsweeps * ns) sequential trials per replica, numbered from
            // sweep0 * ns, in chunks of kMcChunk per launch; the potential
            // starts as the Ewald potential of the state, and each launch
            // also adds the previous chunk's moves to it for the next one
            ewald_field(X, phi1);
            launch_planes_to_double<T>(phi1, mc.phi[0], 3 * (long long)total, st);
            launch_fill_double(mc.acc, 4 * (long long)R, 0.0, st);
            hipEvent_t e = tbegin();
            for (const auto& g : ewg) {  // replicas of one dims at a time
                const uint32_t ns = (uint32_t)reps[g.r0].ncell * (uint32_t)na;
                const long long niters = (long long)std::ceil(nsweeps * (double)ns);
                // several replicas: after every kMcFftPeriod chunks the potential
                // is the FFT field of the state again (one batched evaluation),
                // cheaper than adding every accepted move at every site of every
                // replica
                const bool fft = g.nr >= env_int("SNH_MC_FFT_REPLICAS", kMcFftReplicas);
                const size_t off = reps[g.r0].off;
                const size_t len = (g.r0 + g.nr < R ? (size_t)reps[g.r0 + g.nr].off : (size_t)total) - off;
                int cur = 0;  // mc.phi[cur]: potential before the previous chunk's moves
                for (long long t = 0, k = 0; t < niters; t += kMcChunk, k++) {
                    const uint64_t T0 = sweep0 * ns + (uint64_t)t;
                    const int nt = (int)std::min<long long>(kMcChunk, niters - t);
                    McChunkIO io;
                    io.phi_in = mc.phi[cur];
                    io.phi_out = mc.phi[cur ^ 1];
                    // FFT groups: chunk slot j = k mod kMcFftPeriod sees the moves of
                    // slots 0 .. j-1 on top of the refreshed potential, at its
                    // trial sites only; else the previous chunk's (slot k-1 mod 2)
                    const int slot = fft ? (int)(k % kMcFftPeriod) : (int)(k & 1);
                    const int lists = fft ? slot : (k ? 1 : 0);
                    const int first = fft ? 0 : slot ^ 1;
                    io.prev = lists ? mc.moves + (size_t)first * R * kMcChunk : nullptr;
                    io.nprev = lists ? mc.nmoves + (size_t)first * R : nullptr;
                    io.nprevl = lists;
                    io.nrep_all = R;
                    io.moves = mc.moves + (size_t)slot * R * kMcChunk;
                    io.nmoves = mc.nmoves + (size_t)slot * R;
                    io.pprev = nullptr;
                    if (fft && lists) {
                        // the earlier moves at the chunk's trial sites, summed over
                        // the whole device rather than by the trial workgroups
                        launch_mc_prev_at_trials(A, io, C.reps, g.Aew, na, T0, g.r0, g.nr, mc.pprev, st);
                        io.pprev = mc.pprev;
                    }
                    const uint32_t th = N == 0 ? kMcThreads : 64;  // threads per workgroup of the launch
                    io.napp = k && !fft ? (int)std::min<uint32_t>((ns + th - 1) / th, kMcApplyGroups) : 0;
                    if (N == 0)
                        launch_mc_ewald_chunk<T>(C, A, X, io, g.Aew, g.A0, T0, nt, mc.acc, g.r0, g.nr, st);
                    else
                        launch_sun_mc_ewald_chunk<T>(C, A, X, io, g.Aew, g.A0, T0, nt, mc.acc, g.r0, g.nr, st);
                    if (fft) {
                        if (slot == kMcFftPeriod - 1 && t + kMcChunk < niters) {  // for the next chunks
                            const T* Sg = X;
                            if (N > 0) {
                                launch_sun_expected_spin<T>(C, nblocks, X, Sd, st);
                                Sg = Sd;
                            }
                            // in-LDS transforms write the FP64 potential of the group's
                            // interior sites directly (the trials read no other entry)
                            const bool direct = !g.EG.plain;
                            SNH_REQUIRE(launch_ewald_field<T>(g.EG, g.nr, C.reps + g.r0, total, C.gmat, Sg, g.M,
                                                              g.Fmu, g.Fphi, phi1, g.lib, st,
                                                              direct ? mc.phi[cur] : nullptr),
                                        SNH_ERR_HIP, "hipFFT execution failed");
                            if (!direct)
                                for (int q = 0; q < 3; q++)
                                    launch_planes_to_double<T>(phi1 + q * (size_t)total + off,
                                                               mc.phi[cur] + q * (size_t)total + off, (long long)len,
                                                               st);
                        }
                    } else if (k) {
                        cur ^= 1;
                    }
                }
            }
            tend("local_ewald", e);
            SNH_HIP(hipGetLastError());
            std::vector<double> h(4 * (size_t)R);
            SNH_HIP(hipMemcpyAsync(h.data(), mc.acc, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            if (dE_dS) std::copy(h.begin(), h.end(), dE_dS);
            return;
        }
        launch_fill_double(mc.acc, 4 * (long long)std::max(mc.coff.back(), 1), 0.0, st);
        const long long npass = (long long)std::ceil(2 * nsweeps);
        for (long long k = 0; k < npass; k++) {
            const double prob = k == npass - 1 ? (2 * nsweeps - (double)(npass - 1)) / 2 : 0.5;
            for (int col = 0; col < mc.col.ncolors; col++) {
                hipEvent_t e = tbegin();
                if (N == 0)
                    launch_mc_color<T>(C, A, X, mc.work + mc.coff[col], mc.coff[col + 1] - mc.coff[col],
                                       2 * sweep0 + (uint64_t)k, prob, mc.acc + 4 * (size_t)mc.coff[col], st);
                else
                    launch_sun_mc_color<T>(C, A, X, mc.work + mc.coff[col], mc.coff[col + 1] - mc.coff[col],
                                           2 * sweep0 + (uint64_t)k, prob, mc.acc + 4 * (size_t)mc.coff[col], st);
                tend("local_color", e);
            }
        }
        launch_mc_reduce(mc.acc, mc.start, mc.col.ncolors, R, mc.out, st);
        SNH_HIP(hipGetLastError());
        std::vector<double> h(4 * (size_t)R);
        SNH_HIP(hipMemcpyAsync(h.data(), mc.out, sizeof(double) * h.size(), hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        if (dE_dS) std::copy(h.begin(), h.end(), dE_dS);
    }

    void local_coloring(int32_t* m, int32_t* ncolors, int32_t* color) override
    {
        mc_prepare();
        if (m)
            for (int a = 0; a < 3; a++) m[a] = mc.col.m[a];
        if (ncolors) *ncolors = mc.col.ncolors;
        if (color) std::copy(mc.col.color.begin(), mc.col.color.end(), color);
    }

    void drop_graph()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }

    // One Heun step: Ewald potentials (if any), predictor, corrector.
    void langevin_step(const NoiseArgs<T>& nz, T tdt, T tl, bool dip)
    {
        const T* ewS = nullptr;
        const T* ewSp = nullptr;
        if (ewald) {
            ewald_field(X, phi1, dip ? Dz : nullptr);
            ewS = phi1;
        }
        hipEvent_t e = tbegin();
        if (tiled)
            launch_tile_predict<T>(C, tp.A, tp.threads, X, Xp, Xo, ewS, nz, tdt, tl, st);
        else if (N == 0)
            launch_dipole_predict<T>(C, nblocks, X, Xp, Xo, ewS, nz, tdt, tl, st);
        else
            launch_sun_predict<T>(C, nblocks, X, Xp, Xo, ewS, nz, tdt, tl, dip ? Dz : nullptr, dip ? Dzp : nullptr,
                                  st);
        tend(N == 0 ? "dipole_predict" : "sun_predict", e);
        if (ewald) {
            ewald_field(Xp, phi2, dip ? Dzp : nullptr);
            ewSp = phi2;
        }
        e = tbegin();  // the corrector overwrites the state in place (the predictor consumed it)
        if (tiled)
            launch_tile_correct<T>(C, tc.A, tc.threads, Xp, Xo, X, ewSp, tdt, tl, st);
        else if (N == 0)
            launch_dipole_correct<T>(C, nblocks, Xp, Xo, X, ewSp, tdt, tl, st);
        else
            launch_sun_correct<T>(C, nblocks, Xp, Xo, X, ewSp, tdt, tl, dip ? Dzp : nullptr, dip ? Dz : nullptr, st);
        tend(N == 0 ? "dipole_correct" : "sun_correct", e);
    }

    // Captures graph_steps Langevin steps (noise step counter on the device);
    // on any capture failure graphs stay off for this handle.
    bool capture_langevin(const NoiseArgs<T>& nz, T tdt, T tl, bool dip)
    {
        drop_graph();
        if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            return graph_ok = false;
        }
        for (int k = 0; k < graph_steps; k++) {
            langevin_step(nz, tdt, tl, dip);
            launch_step_advance(d_step, st);
        }
        hipGraph_t g = nullptr;
        if (hipStreamEndCapture(st, &g) != hipSuccess || !g) {
            (void)hipGetLastError();
            return graph_ok = false;
        }
        graph = g;
        if (hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            (void)hipGetLastError();
            drop_graph();
            return graph_ok = false;
        }
        return true;
    }

    void step_langevin(double dt, double damping, const double* kT, const double* noise, uint64_t seed,
                       uint64_t step0, int nsteps) override
    {
        SNH_REQUIRE(std::isfinite(dt) && dt > 0, SNH_ERR_INVALID, "Select positive dt");
        SNH_REQUIRE(damping >= 0, SNH_ERR_INVALID, "Select positive damping");
        SNH_REQUIRE(kT, SNH_ERR_INVALID, "kT array required");
        SNH_REQUIRE(nsteps >= 0, SNH_ERR_INVALID, "nsteps must be non-negative");
        SNH_REQUIRE(!noise || nsteps == 1, SNH_ERR_INVALID, "host-supplied noise covers exactly one step");
        std::vector<T> sc(R);
        bool anynoise = false;
        for (int r = 0; r < R; r++) {
            SNH_REQUIRE(kT[r] >= 0, SNH_ERR_INVALID, "Select nonnegative kT");
            double v = (damping == 0 || kT[r] == 0) ? 0.0 : std::sqrt(2 * dt * damping * kT[r]);
            sc[r] = (T)v;
            anynoise |= v != 0;
        }
        SNH_HIP(hipMemcpyAsync(nscale, sc.data(), sizeof(T) * R, hipMemcpyHostToDevice, st));
        NoiseArgs<T> nz{};
        nz.scale = nscale;
        nz.seed = seed;
        nz.mode = anynoise ? (noise ? 1 : 2) : 0;
        std::vector<double> hn;
        if (nz.mode == 1) {
            if (!host_noise) host_noise = own(dalloc<double>((size_t)planes * total));
            hn.assign((size_t)planes * total, 0.0);
            for_each_padded([&](uint32_t g, long long s, bool interior) {
                if (interior)
                    for (int p = 0; p < planes; p++) hn[(size_t)planes * g + p] = noise[(size_t)planes * s + p];
            });
            SNH_HIP(hipMemcpyAsync(host_noise, hn.data(), sizeof(double) * hn.size(), hipMemcpyHostToDevice, st));
            nz.host = host_noise;
        }
        SNH_HIP(hipStreamSynchronize(st));  // host vectors above are pageable
        const T tdt = (T)dt, tl = (T)damping;
        // SU(N) with bilinear couplings only: the stages exchange expected
        // dipoles (3 planes) instead of kets for the neighbour gather.
        const bool dip = N > 0 && !any_biquad && !tiled;
        if (dip && nsteps > 0) {
            if (!Dz) {
                Dz = own(dalloc<T>(3 * (size_t)total));
                Dzp = own(dalloc<T>(3 * (size_t)total));
            }
            // a run of such steps carries the corrector's dipoles from call to
            // call, so that results do not depend on how steps are split
            if (!dz_valid) launch_sun_expected_spin<T>(C, nblocks, X, Dz, st);
            dz_valid = true;
        }
        if (persist) {
            // chunks bound the length of one launch
            for (int s = 0; s < nsteps; s += 4096) {
                nz.step = step0 + (uint64_t)s;
                hipEvent_t e = tbegin();
                launch_dipole_langevin_persist<T>(C, R, d_rblk, X, Xp, Xo, nz, tdt, tl, std::min(4096, nsteps - s),
                                                  st);
                tend("dipole_persist", e);
            }
            SNH_HIP(hipGetLastError());
            return;
        }
        int s = 0;
        // the first step of a handle runs eagerly (it also sets kernel attributes)
        if (!warmed && nsteps > 0) {
            nz.step = step0;
            langevin_step(nz, tdt, tl, dip);
            SNH_HIP(hipGetLastError());
            warmed = true;
            s = 1;
        }
        // whole chunks of graph_steps steps replay a captured graph (device
        // noise or none, kernel timing off)
        if (graph_ok && graph_steps > 0 && nz.mode != 1 && !timer.enabled && nsteps - s >= graph_steps) {
            if (!d_step) d_step = own(dalloc<unsigned long long>(1));
            const std::array<double, 5> key = {dt, damping, (double)(seed >> 32), (double)(seed & 0xffffffffu),
                                               (double)(nz.mode + 4 * dip)};
            NoiseArgs<T> nzg = nz;
            nzg.step_ptr = d_step;
            if (graph_exec && key != gkey) drop_graph();
            if (!graph_exec && capture_langevin(nzg, tdt, tl, dip)) gkey = key;
            if (graph_exec) {
                launch_step_set(d_step, step0 + (uint64_t)s, st);
                for (; nsteps - s >= graph_steps; s += graph_steps) SNH_HIP(hipGraphLaunch(graph_exec, st));
            }
        }
        if (pair_on && !timer.enabled && s < nsteps) {
            langevin_pairs(nz, tdt, tl, step0, s, nsteps, X);
            s = nsteps;
        }
        for (; s < nsteps; s++) {
            nz.step = step0 + (uint64_t)s;
            langevin_step(nz, tdt, tl, dip);
        }
        SNH_HIP(hipGetLastError());
    }

    void step_midpoint(double dt, double damping, double tol, int max_iters, int nsteps, int* iters) override
    {
        dz_valid = false;
        SNH_REQUIRE(std::isfinite(dt) && dt > 0, SNH_ERR_INVALID, "Select positive dt");
        SNH_REQUIRE(damping >= 0, SNH_ERR_INVALID, "Select nonnegative damping");
        SNH_REQUIRE(max_iters >= 1, SNH_ERR_INVALID, "max_iters must be at least 1");
        ensure_midpoint_buffers();
        // atol = tol * sqrt(length(S)), length counting sites (Integrators.jl:383)
        std::vector<double> a2(R);
        for (int r = 0; r < R; r++) {
            double at = tol * std::sqrt((double)replica_sites(r));
            a2[r] = at * at;
        }
        SNH_HIP(hipMemcpyAsync(atol2, a2.data(), sizeof(double) * R, hipMemcpyHostToDevice, st));
        SNH_HIP(hipStreamSynchronize(st));
        const T tdt = (T)dt, tl = (T)damping;
        if (coop && !tiled && nsteps > 0) {
            // every step iterated to convergence in one cooperative launch
            if (!fail_step) fail_step = own(dalloc<int>(R));
            if (!coop_partial) coop_partial = own(dalloc<double>(2 * (size_t)nblocks));
            if (!coop_bar) coop_bar = own(dalloc<unsigned>(3));
            launch_fill_int(fail_step, R, -1, st);
            launch_fill_int(conv_iter, R, 0, st);
            launch_fill_int(reinterpret_cast<int*>(coop_bar), 3, 0, st);
            const GridBar bar{coop_bar, coop_bar + 1, reinterpret_cast<int*>(coop_bar + 2), 1u << 22};
            hipEvent_t e = tbegin();
            const hipError_t le = launch_dipole_midpoint_coop<T>(C, nblocks, X, B1, B2, Bh, Bh2, tdt, tl, atol2,
                                                                 max_iters, nsteps, conv_iter, fail_step,
                                                                 coop_partial, bar, st);
            tend("midpoint_coop", e);
            if (le != hipSuccess) {
                // a grid the device cannot keep resident: launched sweeps
                // from now on (same layout, same results)
                (void)hipGetLastError();
                coop = false;
                step_midpoint(dt, damping, tol, max_iters, nsteps, iters);
                return;
            }
            SNH_HIP(hipGetLastError());
            int h[3];
            SNH_HIP(hipMemcpyAsync(pinned_i, conv_iter, sizeof(int), hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            h[0] = pinned_i[0];
            SNH_HIP(hipMemcpyAsync(pinned_i, fail_step, sizeof(int), hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            h[1] = pinned_i[0];
            SNH_HIP(hipMemcpyAsync(pinned_i, coop_bar + 2, sizeof(int), hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            h[2] = pinned_i[0];
            SNH_REQUIRE(h[2] == 0, SNH_ERR_HIP, "cooperative midpoint: a grid barrier wait timed out");
            if (iters) iters[0] = h[0];
            last_iters = h[0];
            if (h[1] >= 0) {
                char msg[256];
                snprintf(msg, sizeof msg,
                         "Spherical midpoint method failed to converge to tolerance %g after %d iterations.", tol,
                         max_iters);
                throw Failure{SNH_ERR_NOCONVERGE, msg};
            }
            return;
        }
        if (persist && R == 1 && nsteps > 0) {
            // every step iterated to convergence on the device; one host check per call
            if (!fail_step) fail_step = own(dalloc<int>(R));
            launch_fill_int(fail_step, R, -1, st);
            launch_fill_int(conv_iter, R, 0, st);
            hipEvent_t e = tbegin();
            launch_dipole_midpoint_persist<T>(C, R, d_rblk, X, B1, B2, Bh, Bh2, tdt, tl, atol2, max_iters, nsteps,
                                              conv_iter, fail_step, st);
            tend("midpoint_persist", e);
            SNH_HIP(hipGetLastError());
            SNH_HIP(hipMemcpyAsync(pinned_i, conv_iter, sizeof(int) * R, hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            const int its = pinned_i[0];
            SNH_HIP(hipMemcpyAsync(pinned_i, fail_step, sizeof(int) * R, hipMemcpyDeviceToHost, st));
            SNH_HIP(hipStreamSynchronize(st));
            if (iters) iters[0] = its;
            last_iters = its;
            if (pinned_i[0] >= 0) {
                char msg[256];
                snprintf(msg, sizeof msg,
                         "Spherical midpoint method failed to converge to tolerance %g after %d iterations.", tol,
                         max_iters);
                throw Failure{SNH_ERR_NOCONVERGE, msg};
            }
            return;
        }
        // Groups of steps run optimistically: each step gets a fixed budget of
        // iterations (the last count + 1; iterations after convergence are
        // no-ops) and no host check; a group in which some replica needs more
        // is undone from a checkpoint and rerun with a host check per step.
        // A group that overran its budget is tried once more with two more
        // iterations per step before falling back to per-step checks.
        const bool optimistic = !mid_sync;
        constexpr int kGroup = 16;
        for (int s0 = 0; s0 < nsteps; s0 += kGroup) {
            const int n = std::min(kGroup, nsteps - s0);
            const int budget = std::min(max_iters, last_iters + 1);
            if (optimistic && midpoint_budgeted(n, budget, tdt, tl, iters)) continue;
            if (optimistic && budget < max_iters &&
                midpoint_budgeted(n, std::min(max_iters, budget + 2), tdt, tl, iters))
                continue;
            if (optimistic) ++mid_fallbacks;
            midpoint_checked(n, max_iters, tol, tdt, tl, iters);
        }
    }

    // Fixed-point iteration `it` of a midpoint step: S'' of iteration it goes
    // to bufs[it % 2] (m odd -> B1), then the per-replica convergence check.
    void midpoint_iteration(int it, T tdt, T tl)
    {
        T* bufs[2] = {B2, B1};
        T* S1 = it == 1 ? X : bufs[(it + 1) % 2];  // the first guess S' = S
        T* S2 = bufs[it % 2];
        const T* phi = nullptr;
        const T* hat = nullptr;
        T* hat_next = nullptr;
        if (N == 0) {
            // Shat of iteration it lives in hats[it % 2]: formed from S and
            // S1 = S at it = 1 (unless the previous step's finalize wrote
            // it), then written by the previous iteration
            T* hats[2] = {Bh2, Bh};
            if (it == 1 && !mid_hat_ready) launch_dipole_midpoint_hat<T>(C, nblocks, X, S1, hats[1], conv_iter, st);
            hat = hats[it % 2];
            hat_next = hats[(it + 1) % 2];
            if (ewald) {
                ewald_field(hat, phi1);
                phi = phi1;
            }
        } else if (ewald) {
            launch_sun_midpoint_bar<T>(C, nblocks, X, S1, Bh, conv_iter, st);
            ewald_field(Bh, phi1);
            phi = phi1;
        }
        hipEvent_t e = tbegin();
        if (tiled)  // iteration m keeps mu_m in Bmu + (m odd ? 0 : total)
            launch_tile_midpoint_iter<T>(C, tc.A, tc.threads, X, S1, S2, hat, phi, tdt, tl, partial, conv_iter,
                                         hat_next, it == 1 ? nullptr : Bmu + (it % 2 ? total : 0),
                                         Bmu + (it % 2 ? 0 : total), st);
        else if (N == 0)
            launch_dipole_midpoint_iter<T>(C, nblocks, X, S1, S2, hat, phi, tdt, tl, partial, conv_iter, hat_next,
                                           st);
        else
            launch_sun_midpoint_iter<T>(C, nblocks, X, S1, S2, phi, tdt, tl, partial, conv_iter, st);
        tend("midpoint_iter", e);
        launch_midpoint_check(partial, tiled ? d_tcblk : d_rblk, R, atol2, conv_iter, it, st);
    }

    // S = normalize(S'') for every converged replica; with next_hat (:dipole)
    // also the first Shat of the next step.
    void midpoint_finalize(bool next_hat = false)
    {
        if (N == 0 && tiled)  // Shat of iteration m + 1 in hats[(m + 1) % 2] = {Bh2, Bh}
            launch_dipole_midpoint_finalize_mu<T>(C, nblocks, X, Bh2, Bh, Bmu, Bmu + total, conv_iter,
                                                  next_hat ? Bh : nullptr, st);
        else if (N == 0)
            launch_dipole_midpoint_finalize<T>(C, nblocks, X, B1, B2, conv_iter, 0, next_hat ? Bh : nullptr, st);
        else
            launch_sun_midpoint_finalize<T>(C, nblocks, X, B1, B2, conv_iter, 0, st);
    }

    // n steps of `budget` iterations each, with one host check at the end.
    // Returns false, with the state restored to the start, when a replica
    // had not converged within the budget at some step.
    bool midpoint_budgeted(int n, int budget, T tdt, T tl, int* iters)
    {
        if (!fail_step) fail_step = own(dalloc<int>(R));
        if (!max_iter) max_iter = own(dalloc<int>(R));
        const size_t bytes = sizeof(T) * planes * (size_t)total;
        launch_copy_planes(X, Xo, bytes, st);  // checkpoint (Xo is Heun scratch)
        launch_fill_int(fail_step, R, -1, st);
        launch_fill_int(conv_iter, R, 0, st);  // k_mid_budget clears it after every step
        for (int s = 0; s < n; s++) {
            // after the first step, the previous finalize wrote this step's first Shat
            mid_hat_ready = s > 0;
            for (int it = 1; it <= budget; it++) midpoint_iteration(it, tdt, tl);
            midpoint_finalize(s + 1 < n);
            launch_mid_budget(conv_iter, fail_step, max_iter, last_it, R, s, st);
        }
        mid_hat_ready = false;
        SNH_HIP(hipGetLastError());
        SNH_HIP(hipMemcpyAsync(pinned_i, fail_step, sizeof(int) * R, hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        bool ok = true;
        for (int r = 0; r < R; r++) ok &= pinned_i[r] < 0;
        if (!ok) {
            ++mid_reruns;
            launch_copy_planes(Xo, X, bytes, st);
            SNH_HIP(hipGetLastError());
            return false;
        }
        // the next group's budget covers the slowest step of this one
        SNH_HIP(hipMemcpyAsync(pinned_i, max_iter, sizeof(int) * R, hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        int mx = 0;
        for (int r = 0; r < R; r++) mx = std::max(mx, pinned_i[r]);
        last_iters = mx;
        SNH_HIP(hipMemcpyAsync(pinned_i, last_it, sizeof(int) * R, hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        if (iters)
            for (int r = 0; r < R; r++) iters[r] = pinned_i[r];
        return true;
    }

    // n steps, each iterated with host convergence checks until every replica
    // converged (or max_iters: the reference's error).
    void midpoint_checked(int n, int max_iters, double tol, T tdt, T tl, int* iters)
    {
        for (int s = 0; s < n; s++) {
            launch_fill_int(conv_iter, R, 0, st);
            int it = 0;
            bool done = false;
            int chunk = std::max(1, std::min(last_iters, max_iters));
            while (!done) {
                const int upto = std::min(max_iters, it + chunk);
                while (it < upto) midpoint_iteration(++it, tdt, tl);
                SNH_HIP(hipGetLastError());
                SNH_HIP(hipMemcpyAsync(pinned_i, conv_iter, sizeof(int) * R, hipMemcpyDeviceToHost, st));
                SNH_HIP(hipStreamSynchronize(st));
                bool all = true;
                for (int r = 0; r < R; r++) all &= pinned_i[r] != 0;
                // S = normalize(S'') for every converged replica, once per step
                // (also before reporting a replica that did not converge)
                if (all || it >= max_iters) {
                    midpoint_finalize();
                    SNH_HIP(hipGetLastError());
                }
                if (all) {
                    done = true;
                } else if (it >= max_iters) {
                    char msg[256];
                    if (N == 0)
                        snprintf(msg, sizeof msg,
                                 "Spherical midpoint method failed to converge to tolerance %g after %d iterations.",
                                 tol, max_iters);
                    else
                        snprintf(msg, sizeof msg, "Schrödinger midpoint method failed to converge in %d iterations.",
                                 max_iters);
                    throw Failure{SNH_ERR_NOCONVERGE, msg};
                } else {
                    chunk = 1;
                }
            }
            int mx = 0;
            for (int r = 0; r < R; r++) mx = std::max(mx, pinned_i[r]);
            last_iters = mx;
            if (iters)
                for (int r = 0; r < R; r++) iters[r] = pinned_i[r];
        }
    }

    // ---- adjoint of the deterministic midpoint trajectory

    void require_adjoint_support() const
    {
        SNH_REQUIRE(N == 0, SNH_ERR_UNSUPPORTED, "the midpoint adjoint supports :dipole mode");
        SNH_REQUIRE(sizeof(T) == 8, SNH_ERR_UNSUPPORTED, "the midpoint adjoint runs at precision 64");
        SNH_REQUIRE(R == 1, SNH_ERR_UNSUPPORTED, "the midpoint adjoint takes one replica");
        SNH_REQUIRE(!tiled, SNH_ERR_UNSUPPORTED, "the midpoint adjoint needs the ghost-padded layout");
    }

    // One midpoint step (damping 0); returns the buffer holding S'' of the
    // converged iteration and its iteration count.
    T* midpoint_one(double dt, double tol, int max_iters, int* it_out)
    {
        step_midpoint(dt, 0.0, tol, max_iters, 1, it_out);
        return (*it_out % 2 == 1) ? B1 : B2;  // S'' of iteration m lives in B1 for odd m, B2 for even m
    }

    void midpoint_record(double dt, double tol, int max_iters, int nsteps, int every) override
    {
        dz_valid = false;
        require_adjoint_support();
        SNH_REQUIRE(nsteps >= 1 && every >= 1, SNH_ERR_INVALID, "nsteps and checkpoint_every must be positive");
        const int ncp = (nsteps + every - 1) / every;
        const size_t bytes = sizeof(T) * planes * (size_t)total;
        for (T* p : ckpt) (void)hipFree(p);
        ckpt.clear();
        rec_n = 0;
        for (int c = 0; c < ncp; c++) ckpt.push_back(dalloc<T>((size_t)planes * total));
        rec_iters.assign(nsteps, 0);
        rec_every = every;
        rec_dt = dt;
        rec_tol = tol;
        rec_maxit = max_iters;
        for (int n = 0; n < nsteps; n++) {
            if (n % every == 0) SNH_HIP(hipMemcpyAsync(ckpt[n / every], X, bytes, hipMemcpyDeviceToDevice, st));
            int it = 0;
            midpoint_one(dt, tol, max_iters, &it);
            rec_iters[n] = it;
        }
        SNH_HIP(hipStreamSynchronize(st));
        rec_n = nsteps;
    }

    // Unit-coupling tables of one parameter, laid out like C (same bond slots).
    Couplings<T> param_tables(const snh_spec& ps, const T* zeros3)
    {
        SNH_REQUIRE(ps.na == na && ps.N == 0, SNH_ERR_INVALID, "parameter spec: na and mode must match the system");
        SNH_REQUIRE(ps.nbonds == nb_in, SNH_ERR_INVALID, "parameter spec: bond list must match the system's");
        for (int k = 0; k < 6 * nb_in; k++)
            if (k % 6 != 5)
                SNH_REQUIRE(ps.bond[k] == bond_h[k], SNH_ERR_INVALID,
                            "parameter spec: bond list must match the system's");
        const size_t nbp = bsrc_h.size();
        std::vector<int> flag(nbp, kBondIso | kBondCulled);
        std::vector<T> J(9 * nbp, (T)0), J0(nbp, (T)0), Q(25 * nbp, (T)0), q2(9 * na, (T)0), stev(28 * na, (T)0);
        std::vector<int> km(na, 0);
        for (size_t k = 0; k < nbp; k++) {
            const int b = bsrc_h[k];
            if (b < 0) continue;
            const double* Jb = ps.bilin + 9 * b;
            const bool iso = Jb[1] == 0 && Jb[2] == 0 && Jb[3] == 0 && Jb[5] == 0 && Jb[6] == 0 && Jb[7] == 0 &&
                             Jb[0] == Jb[4] && Jb[0] == Jb[8];
            const bool bq = ps.has_biquad && ps.has_biquad[b];
            flag[k] = (iso ? kBondIso : 0) | (bq ? kBondBiquad : 0);
            for (int p = 0; p < 9; p++) J[9 * k + p] = (T)Jb[p];
            J0[k] = (T)Jb[0];
            if (bq)
                for (int p = 0; p < 25; p++) Q[25 * k + p] = (T)ps.biquad[25 * b + p];
        }
        for (int i = 0; i < na; i++) {
            const int k = ps.kmax ? ps.kmax[i] : 0;
            SNH_REQUIRE(k == 0 || k == 2 || k == 4 || k == 6, SNH_ERR_INVALID, "parameter spec: kmax must be 0, 2, 4 or 6");
            km[i] = k;
            if (k > 0)
                for (int p = 0; p < 28; p++) stev[28 * i + p] = (T)ps.stevens[28 * i + p];
            if (k == 2) {
                const double* c2 = ps.stevens + 28 * i + 1;
                const double Q2[9] = {2 * c2[0], 2 * c2[4], c2[1], 2 * c2[4], -2 * c2[0], c2[3], c2[1], c2[3], 6 * c2[2]};
                for (int p = 0; p < 9; p++) q2[9 * i + p] = (T)Q2[p];
            }
        }
        Couplings<T> Cp = C;
        Cp.bflag = own(dupload(flag));
        Cp.bJ = own(dupload(J));
        Cp.bJ0 = own(dupload(J0));
        Cp.bQ = own(dupload(Q));
        Cp.stev = own(dupload(stev));
        Cp.stev_q2 = own(dupload(q2));
        Cp.kmax = own(dupload(km));
        Cp.zee = zeros3;
        Cp.iso = 0;
        return Cp;
    }

    void midpoint_adjoint(int nparams, const snh_spec* params, const double* dLdS, double* grad,
                          double* dLdS0) override
    {
        if constexpr (sizeof(T) != 8) {
            throw Failure{SNH_ERR_UNSUPPORTED, "the midpoint adjoint runs at precision 64"};
        } else {
            require_adjoint_support();
            SNH_REQUIRE(rec_n > 0, SNH_ERR_INVALID, "no recorded trajectory: call snh_midpoint_record first");
            SNH_REQUIRE(nparams >= 0 && (nparams == 0 || (params && grad)), SNH_ERR_INVALID, "invalid parameters");
            SNH_REQUIRE(dLdS, SNH_ERR_INVALID, "null loss gradient");
            const size_t np = (size_t)planes * total, bytes = sizeof(T) * np;
            // scratch for this call (freed below)
            std::vector<void*> tmp;
            auto alloc = [&](size_t n) {
                T* p = dalloc<T>(n);
                tmp.push_back(p);
                return p;
            };
            struct Guard {
                std::vector<void*>& v;
                ~Guard()
                {
                    for (void* p : v) (void)hipFree(p);
                }
            } guard{tmp};
            const size_t owned0 = owned.size();
            std::vector<T> z3(3 * (size_t)na, (T)0);
            T* zeros3 = dalloc<T>(z3.size());
            tmp.push_back(zeros3);
            SNH_HIP(hipMemcpy(zeros3, z3.data(), sizeof(T) * z3.size(), hipMemcpyHostToDevice));
            std::vector<Couplings<T>> Cps;
            for (int p = 0; p < nparams; p++) Cps.push_back(param_tables(params[p], zeros3));
            // param_tables' uploads belong to this call too
            for (size_t k = owned0; k < owned.size(); k++) tmp.push_back(owned[k]);
            owned.resize(owned0);
            Couplings<T> Clin = C;
            Clin.zee = zeros3;
            Clin.iso = 0;
            // the nonlinear terms leave Clin: adj_nonlinear_jvp differentiates them at Shat
            bool nonlinear = any_biquad;
            for (int k : kmax_h) nonlinear |= k > 2;
            if (nonlinear) {
                std::vector<int> fl(bflag_h), km(kmax_h);
                for (int& f : fl) f &= ~kBondBiquad;
                for (int& k : km) k = k > 2 ? 0 : k;
                int* dfl = dupload(fl);
                tmp.push_back(dfl);
                int* dkm = dupload(km);
                tmp.push_back(dkm);
                Clin.bflag = dfl;
                Clin.kmax = dkm;
            }

            T *Xsave = alloc(np), *lam = alloc(np), *mu = alloc(np), *nu = alloc(np), *nu2 = alloc(np),
              *Shat = alloc(np), *w = alloc(np), *g = alloc(np);
            double* dgrad = (double*)dalloc<double>(std::max(1, nparams));
            tmp.push_back(dgrad);
            const int K = rec_every;
            std::vector<T*> segS, segD;
            for (int k = 0; k < K; k++) {
                segS.push_back(alloc(np));
                segD.push_back(alloc(np));
            }
            SNH_HIP(hipMemcpyAsync(Xsave, X, bytes, hipMemcpyDeviceToDevice, st));
            SNH_HIP(hipMemsetAsync(dgrad, 0, sizeof(double) * std::max(1, nparams), st));
            SNH_HIP(hipMemsetAsync(lam, 0, bytes, st));
            {
                std::vector<T> h(np, (T)0);
                for_each_padded([&](uint32_t gi, long long si, bool interior) {
                    if (!interior) return;
                    for (int a = 0; a < 3; a++) h[(size_t)a * total + gi] = (T)dLdS[3 * (size_t)si + a];
                });
                SNH_HIP(hipMemcpyAsync(lam, h.data(), bytes, hipMemcpyHostToDevice, st));
                SNH_HIP(hipStreamSynchronize(st));
            }
            const T tdt = (T)rec_dt;
            const int ncp = (int)ckpt.size();
            for (int seg = ncp - 1; seg >= 0; seg--) {
                const int n0 = seg * K, n1 = std::min(n0 + K, rec_n);
                SNH_HIP(hipMemcpyAsync(X, ckpt[seg], bytes, hipMemcpyDeviceToDevice, st));
                for (int n = n0; n < n1; n++) {
                    SNH_HIP(hipMemcpyAsync(segS[n - n0], X, bytes, hipMemcpyDeviceToDevice, st));
                    int it = 0;
                    T* Sdd = midpoint_one(rec_dt, rec_tol, rec_maxit, &it);
                    SNH_HIP(hipMemcpyAsync(segD[n - n0], Sdd, bytes, hipMemcpyDeviceToDevice, st));
                    rec_iters[n] = std::max(rec_iters[n], it);
                }
                for (int n = n1 - 1; n >= n0; n--) {
                    const T* S = segS[n - n0];
                    const T* Sdd = segD[n - n0];
                    launch_adj_prep<T>(C, nblocks, S, Sdd, lam, Shat, mu, st);
                    // the field at Shat, with its Ewald potential (phi1)
                    if (ewald) ewald_field(Shat, phi1);
                    launch_dipole_grad<T>(C, nblocks, Shat, ewald ? phi1 : nullptr, g, st);
                    SNH_HIP(hipMemcpyAsync(nu, mu, bytes, hipMemcpyDeviceToDevice, st));
                    // the linear sweep contracts like the forward fixed point
                    const int sweeps = rec_iters[n] + 2;
                    for (int k = 0; k < sweeps; k++) {
                        launch_adj_w<T>(C, nblocks, nu, Shat, tdt, w, st);
                        // the Ewald part of M w: the field of w's Ewald potential (phi2)
                        if (ewald) ewald_field(w, phi2);
                        launch_adj_apply<T>(Clin, nblocks, w, S, Sdd, g, nu, mu, tdt, nu2, partial,
                                            ewald ? phi2 : nullptr, C, nonlinear ? Shat : nullptr, st);
                        std::swap(nu, nu2);
                    }
                    launch_adj_w<T>(C, nblocks, nu, Shat, tdt, w, st);
                    for (int p = 0; p < nparams; p++) {
                        launch_adj_param<T>(Cps[p], nblocks, Shat, w, partial, st);
                        launch_adj_accumulate(partial, nblocks, dgrad, p, st);
                    }
                    launch_adj_lambda<T>(C, nblocks, nu, mu, lam, st);
                }
                SNH_HIP(hipGetLastError());
            }
            SNH_HIP(hipMemcpyAsync(X, Xsave, bytes, hipMemcpyDeviceToDevice, st));
            if (nparams > 0) {
                SNH_HIP(hipMemcpyAsync(grad, dgrad, sizeof(double) * nparams, hipMemcpyDeviceToHost, st));
            }
            SNH_HIP(hipStreamSynchronize(st));
            if (dLdS0) download_planes(lam, 3, dLdS0);
        }
    }

    // Queues energy_aux of every replica on the handle's stream: per-workgroup
    // partial sums into `part` (one entry per workgroup of the energy launch,
    // each written by its workgroup), then their sums per replica into out[R].
    void energy_launch(double* part, double* out, bool timed)
    {
        const T* phi = nullptr;
        if (ewald) {
            ewald_field(X, phi1, nullptr, timed);
            phi = phi1;
        }
        if (tiled)
            launch_tile_energy<T>(C, tp.A, tp.threads, X, phi, part, st);
        else if (N == 0)
            launch_dipole_energy<T>(C, nblocks, X, phi, part, st);
        else
            launch_sun_energy<T>(C, nblocks, X, phi, part, st);
        launch_reduce_partials(part, tiled ? d_trblk : d_rblk, R, out, st);
        SNH_HIP(hipGetLastError());
    }

    void energy(double* E) override
    {
        energy_launch(partial, d_out, true);
        SNH_HIP(hipMemcpyAsync(pinned, d_out, sizeof(double) * R, hipMemcpyDeviceToHost, st));
        SNH_HIP(hipStreamSynchronize(st));
        for (int r = 0; r < R; r++) E[r] = pinned[r];
    }

    // energy() into device memory with the partial sums in the static scratch:
    // no allocation, no host synchronisation, no event.
    void energy_noalloc(double* E_dev) override
    {
        SNH_REQUIRE(E_dev, SNH_ERR_INVALID, "null energy buffer");
        const size_t need = tiled ? (size_t)tp.A.nwork : (size_t)nblocks;
        SNH_REQUIRE(need <= kNoAllocPartials, SNH_ERR_UNSUPPORTED,
                    "snh_energy_noalloc: the energy launch needs " + std::to_string(need) +
                        " partial sums, more than the static scratch holds (" + std::to_string(kNoAllocPartials) +
                        "); use snh_energy");
        SNH_REQUIRE(!(ewald && ewald_lib_any()), SNH_ERR_UNSUPPORTED,
                    "snh_energy_noalloc: the Ewald grid is transformed by hipFFT; use snh_energy");
        double* part = noalloc_scratch(dev);
        SNH_REQUIRE(part, SNH_ERR_HIP, "static scratch not resolved on this device");
        energy_launch(part, E_dev, false);
    }

    void energy_grad(double* out) override
    {
        const T* phi = nullptr;
        if (ewald) {
            ewald_field(X, phi1);
            phi = phi1;
        }
        if (tiled)
            launch_tile_grad<T>(C, tp.A, tp.threads, X, phi, Xp, st);
        else if (N == 0)
            launch_dipole_grad<T>(C, nblocks, X, phi, Xp, st);
        else
            launch_sun_grad<T>(C, nblocks, X, phi, Xp, st);
        SNH_HIP(hipGetLastError());
        download_planes(Xp, planes, out);
    }

    void synchronize() override { SNH_HIP(hipStreamSynchronize(st)); }

    void enable_timing(bool on) override
    {
        SNH_HIP(hipStreamSynchronize(st));
        timer.clear();
        timer.enabled = on;
    }

    void copy_calibration(int repeats, int64_t* bytes) override
    {
        const size_t b = (sizeof(T) * planes * (size_t)total) / 16 * 16;
        for (int k = 0; k < repeats; k++) {
            hipEvent_t e = tbegin();
            launch_copy_planes(X, Xp, b, st);
            tend("copy_planes", e);
        }
        SNH_HIP(hipGetLastError());
        SNH_HIP(hipStreamSynchronize(st));
        *bytes = (int64_t)(2 * b);
    }

    int64_t query(const std::string& key) const override
    {
        if (key == "layout") return tiled ? SNH_LAYOUT_AUTO : SNH_LAYOUT_GHOST;
        if (key == "tile_T1") return tiled ? tp.A.T1 : 0;
        if (key == "tile_T2") return tiled ? tp.A.T2 : 0;
        if (key == "tile_threads") return tiled ? tp.threads : 0;
        if (key == "tile_sp") return tiled ? tp.A.sp : 0;
        if (key == "tile_work") return tiled ? tp.A.nwork : 0;
        if (key == "tile_lds") return tiled ? (int64_t)tp.lds : 0;
        if (key == "tile_candidates") return (int64_t)tile_cands.size();
        if (key == "tile_tuned") return tuned_index[0];
        if (key == "tile_tuned_correct") return tuned_index[1];
        if (key == "correct_T1") return tiled ? tc.A.T1 : 0;
        if (key == "correct_T2") return tiled ? tc.A.T2 : 0;
        if (key == "correct_sp") return tiled ? tc.A.sp : 0;
        if (key == "correct_threads") return tiled ? tc.threads : 0;
        if (key == "correct_work") return tiled ? tc.A.nwork : 0;
        if (key == "gather_chunk") return C.chunk;
        if (key == "graph_steps") return graph_steps;
        if (key == "persistent") return persist ? 1 : 0;
        if (key == "cooperative") return coop ? 1 : 0;
        if (key == "graph_live") return graph_exec ? 1 : 0;
        if (key == "pair_split") return pair_split;
        if (key == "pair_on") return pair_on ? 1 : 0;
        if (key == "mid_reruns") return mid_reruns;
        if (key == "mid_fallbacks") return mid_fallbacks;
        if (key == "ewald_fft_lib") return ewald && ewald_lib_any() ? 1 : 0;
        if (key == "ewald_slab") {
            int n = 0;
            for (const auto& g : ewg) n += g.EG.slab;
            return n;  // groups on the slab path
        }
        if (key == "corr_fft_lib") return corr.lib ? 1 : 0;
        throw Failure{SNH_ERR_INVALID, "unknown query key: " + key};
    }

    void kernel_time(const char* fam, double* ms, int64_t* n) override
    {
        SNH_HIP(hipStreamSynchronize(st));
        double acc = 0;
        int64_t cnt = 0;
        for (auto& ev : timer.events) {
            if (ev.first != fam) continue;
            float t = 0;
            SNH_HIP(hipEventElapsedTime(&t, ev.second.first, ev.second.second));
            acc += t;
            cnt++;
        }
        *ms = acc;
        *n = cnt;
    }
};

int fail(snh_system* sys, const Failure& f)
{
    if (sys) sys->err = f.msg;
    return f.code;
}

template <typename F>
int guarded(snh_system* sys, F&& fn)
{
    if (!sys) return SNH_ERR_INVALID;
    try {
        fn();
        return SNH_OK;
    } catch (const Failure& f) {
        return fail(sys, f);
    } catch (const std::exception& e) {
        return fail(sys, Failure{SNH_ERR_INVALID, e.what()});
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// extern "C" entry points

extern "C" int snh_create(const snh_spec* spec, snh_system** out)
{
    if (!spec || !out) {
        g_create_error = "null argument";
        return SNH_ERR_INVALID;
    }
    *out = nullptr;
    snh_system* sys = nullptr;
    try {
        SNH_REQUIRE(spec->na >= 1, SNH_ERR_INVALID, "na must be at least 1");
        SNH_REQUIRE(spec->N == 0 || (spec->N >= 2 && spec->N <= 7), SNH_ERR_UNSUPPORTED,
                    "supported modes: :dipole (N = 0) and :SUN with 2 <= N <= 7");
        SNH_REQUIRE(spec->precision == 32 || spec->precision == 64, SNH_ERR_INVALID, "precision must be 32 or 64");
        SNH_REQUIRE(spec->nreplicas >= 1 && spec->dims, SNH_ERR_INVALID, "need at least one replica");
        SNH_REQUIRE(spec->kappa && spec->g && spec->field, SNH_ERR_INVALID, "kappa, g and field are required");
        SNH_REQUIRE(spec->nbonds == 0 || (spec->bond && spec->bilin), SNH_ERR_INVALID, "bond tables missing");
        int ndev = 0;
        SNH_HIP(hipGetDeviceCount(&ndev));
        SNH_REQUIRE(ndev > 0, SNH_ERR_HIP, "no HIP device");
        if (spec->precision == 64) {
            auto* s = new SystemImpl<double>();
            sys = s;
            s->build(*spec);
        } else {
            auto* s = new SystemImpl<float>();
            sys = s;
            s->build(*spec);
        }
        *out = sys;
        return SNH_OK;
    } catch (const Failure& f) {
        g_create_error = f.msg;
        delete sys;
        return f.code;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        delete sys;
        return SNH_ERR_INVALID;
    }
}

extern "C" int snh_energy_noalloc(snh_system* sys, double* E_dev)
{
    return guarded(sys, [&] { sys->energy_noalloc(E_dev); });
}

extern "C" int snh_destroy(snh_system* sys)
{
    delete sys;
    return SNH_OK;
}

extern "C" const char* snh_last_error(const snh_system* sys)
{
    return sys ? sys->err.c_str() : g_create_error.c_str();
}

#ifndef SNH_BUILD_ID
#define SNH_BUILD_ID "unknown"
#endif

extern "C" const char* snh_build_id(void) { return SNH_BUILD_ID; }

extern "C" int snh_num_sites(const snh_system* sys, int64_t* total)
{
    if (!sys || !total) return SNH_ERR_INVALID;
    *total = sys->total_sites();
    return SNH_OK;
}

extern "C" int snh_replica_sites(const snh_system* sys, int32_t r, int64_t* n)
{
    if (!sys || !n) return SNH_ERR_INVALID;
    auto* s = const_cast<snh_system*>(sys);
    return guarded(s, [&] {
        SNH_REQUIRE(r >= 0 && r < sys->replicas(), SNH_ERR_INVALID, "replica index out of range");
        *n = sys->replica_sites(r);
    });
}

extern "C" int snh_set_dipoles(snh_system* sys, const double* x)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(x, SNH_ERR_INVALID, "null state");
        sys->set_state(x, false);
    });
}

extern "C" int snh_get_dipoles(snh_system* sys, double* x)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(x, SNH_ERR_INVALID, "null state");
        sys->get_state(x, false);
    });
}

extern "C" int snh_set_coherents(snh_system* sys, const double* x)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(x, SNH_ERR_INVALID, "null state");
        sys->set_state(x, true);
    });
}

extern "C" int snh_get_coherents(snh_system* sys, double* x)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(x, SNH_ERR_INVALID, "null state");
        sys->get_state(x, true);
    });
}

extern "C" int snh_step_langevin(snh_system* sys, double dt, double damping, const double* kT, const double* noise,
                                 uint64_t seed, uint64_t step0, int32_t nsteps)
{
    return guarded(sys, [&] { sys->step_langevin(dt, damping, kT, noise, seed, step0, nsteps); });
}

extern "C" int snh_step_midpoint(snh_system* sys, double dt, double damping, double tol, int32_t max_iters,
                                 int32_t nsteps, int32_t* iters_out)
{
    return guarded(sys, [&] { sys->step_midpoint(dt, damping, tol, max_iters, nsteps, iters_out); });
}

extern "C" int snh_energy(snh_system* sys, double* E)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(E, SNH_ERR_INVALID, "null output");
        sys->energy(E);
    });
}

extern "C" int snh_energy_grad(snh_system* sys, double* out)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(out, SNH_ERR_INVALID, "null output");
        sys->energy_grad(out);
    });
}

extern "C" int snh_synchronize(snh_system* sys)
{
    return guarded(sys, [&] { sys->synchronize(); });
}

extern "C" int snh_enable_timing(snh_system* sys, int32_t enable)
{
    return guarded(sys, [&] { sys->enable_timing(enable != 0); });
}

extern "C" int snh_kernel_time(snh_system* sys, const char* fam, double* ms, int64_t* n)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(fam && ms && n, SNH_ERR_INVALID, "null argument");
        sys->kernel_time(fam, ms, n);
    });
}

extern "C" int snh_copy_calibration(snh_system* sys, int32_t repeats, int64_t* bytes)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(bytes && repeats >= 0, SNH_ERR_INVALID, "invalid argument");
        sys->copy_calibration(repeats, bytes);
    });
}

extern "C" int snh_query(snh_system* sys, const char* key, int64_t* value)
{
    return guarded(sys, [&] {
        SNH_REQUIRE(key && value, SNH_ERR_INVALID, "null argument");
        *value = sys->query(key);
    });
}

extern "C" int snh_corr_init(snh_system* sys, int32_t nobs, const double* obs, int32_t ncorr, const int32_t* pairs,
                             int32_t n_all_omega, int32_t measperiod, int32_t window)
{
    return guarded(sys, [&] { sys->corr_init(nobs, obs, ncorr, pairs, n_all_omega, measperiod, window); });
}

extern "C" int snh_corr_add_sample(snh_system* sys, double dt, double tol, int32_t max_iters)
{
    return guarded(sys, [&] { sys->corr_add_sample(dt, tol, max_iters); });
}

extern "C" int snh_corr_data(snh_system* sys, double* data, int64_t* nsamples)
{
    return guarded(sys, [&] { sys->corr_data(data, nsamples); });
}

extern "C" int snh_step_local(snh_system* sys, const double* kT, int32_t nprop, const int32_t* kinds,
                              const double* magnitudes, const double* weights, uint64_t seed, uint64_t sweep0,
                              double nsweeps, double* dE_dS)
{
    return guarded(sys, [&] { sys->step_local(kT, nprop, kinds, magnitudes, weights, seed, sweep0, nsweeps, dE_dS); });
}

extern "C" int snh_local_coloring(snh_system* sys, int32_t* m, int32_t* ncolors, int32_t* color)
{
    return guarded(sys, [&] { sys->local_coloring(m, ncolors, color); });
}

extern "C" int snh_autotune(snh_system* sys, int32_t candidates, int32_t launches, double* best_ms)
{
    return guarded(sys, [&] { sys->autotune(candidates, launches, best_ms); });
}

extern "C" int snh_midpoint_record(snh_system* sys, double dt, double tol, int32_t max_iters, int32_t nsteps,
                                   int32_t checkpoint_every)
{
    return guarded(sys, [&] { sys->midpoint_record(dt, tol, max_iters, nsteps, checkpoint_every); });
}

extern "C" int snh_midpoint_adjoint(snh_system* sys, int32_t nparams, const snh_spec* params,
                                    const double* dLdS_final, double* grad_out, double* dLdS_initial)
{
    return guarded(sys, [&] { sys->midpoint_adjoint(nparams, params, dLdS_final, grad_out, dLdS_initial); });
}
