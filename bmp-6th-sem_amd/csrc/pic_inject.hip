// pic_inject.hip — new particles after the push (the reference's
// generate_new_electrons and generate_pairs, utils.cpp:374-499,610-777, called
// after the movers in main.cpp:232-245):
//   * re-injected electrons: as many as the electrons that left through x < 0
//     since the last injection minus the ions that did (never negative), at
//     x = x_inject, y uniform in [0, height), Maxwellian;
//   * electron-ion pairs: target = ceil(H dt (2/PI) (x22 - x11) ionimax / 1.67e6)
//     per step; attempt a draws x = xm + asin(2u - 1) (x22 - x11) / PI, y
//     uniform, and is accepted when 0 < x < width; the first `target` accepted
//     attempts (in attempt order, as the reference's sequential loop) give an
//     electron and an ion at the same position.
// Draws come from a counter-based generator: uniform u(seed, injection,
// stream, index, draw) from two SplitMix64 finalizer rounds, normals by the
// Box-Muller transform; the reference's time-seeded std::rand / ran2 give the
// same distributions but no reproducible sequence.
//
// The new particles are staged in storage layout with their cell keys and,
// inside pic_step, counted into the same cell counts as the pushed particles
// before the prefix scan, then scattered after them: the step keeps one sort.
// Staged counts are clamped so the storage never exceeds its capacity.
#include "pic_particle_dev.h"

#include <algorithm>
#include <cmath>

namespace pic {

// device counters (long long each)
enum {
    IS_STEP = 0,   // injections so far (the generator's step index)
    IS_CONS_E,     // x < 0 exit counters of species 0 / 1 at the last injection
    IS_CONS_I,
    IS_RE,         // re-injected electrons staged by the current injection
    IS_NP,         // pairs staged by the current injection
    IS_NE,         // particles added to species 0 / 1 by the current injection
    IS_NI,
    IS_LIVE_E,     // live particles of species 0 / 1 before the current injection
    IS_LIVE_I,
    IS_TOT_RE,     // totals: re-injected electrons, pairs, dropped particles
    IS_TOT_NP,
    IS_TOT_DROP,
    IS_DROP_NOW,   // dropped by the current injection so far
    IS_N
};

static constexpr double kPi = 3.141592654;   // the reference's PI (utils.cpp:17)
static constexpr int kPairBlock = 1024;

struct InjectDev {
    double x_inj, sig_re, sig_e, sig_i;   // plane; velocity standard deviations sqrt(|q/m| T)
    double x11, x22, xm, width, height;
    int reinject, target, attempts, ecap;
    unsigned long long seed;
};

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in (0, 1): 52 random bits + 1/2, times 2^-52
__device__ __forceinline__ double urand(unsigned long long seed, long long step, int stream, long long k, int j) {
    unsigned long long z = mix64(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(step + 1));
    z = mix64(z ^ ((unsigned long long)stream << 60) ^ ((unsigned long long)k << 8) ^ (unsigned long long)j);
    return ((double)(z >> 12) + 0.5) * 0x1p-52;
}

// Box-Muller pair from draws j, j + 1
__device__ __forceinline__ void normal2(unsigned long long seed, long long step, int stream, long long k, int j,
                                        double& a, double& b) {
    const double u = urand(seed, step, stream, k, j), v = urand(seed, step, stream, k, j + 1);
    const double r = sqrt(-2.0 * log(u));
    const double t = 6.283185307179586 * v;
    a = r * cos(t);
    b = r * sin(t);
}

// velocities from draws j .. j + 3
__device__ __forceinline__ void maxwell(unsigned long long seed, long long step, int stream, long long k, int j,
                                        double sig, double& vx, double& vy, double& vz) {
    double a, b, c, d;
    normal2(seed, step, stream, k, j, a, b);
    normal2(seed, step, stream, k, j + 2, c, d);
    vx = sig * a;
    vy = sig * b;
    vz = sig * c;
}

// absolute position -> storage layout + key (the upload encoding)
template <int L>
__device__ __forceinline__ void put(const ParticleBuf& b, int32_t* key, int i, double px, double py, double vx,
                                    double vy, double vz, const GridDev& g) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    const double tx = px * g.inv_dx, ty = py * g.inv_dy;
    const int gx = (int)tx, gy = (int)ty;
    const int k = gy * g.gx + gx;
    if (L == PIC_LAYOUT_F64) {
        ((P*)b.pa)[i] = (P)px;
        ((P*)b.pb)[i] = (P)py;
    } else {
        ((P*)b.pa)[i] = clamp_offset((float)(tx - (double)gx));
        ((P*)b.pb)[i] = clamp_offset((float)(ty - (double)gy));
    }
    ((V*)b.vx)[i] = (V)vx;
    ((V*)b.vy)[i] = (V)vy;
    ((V*)b.vz)[i] = (V)vz;
    b.cell[i] = k;
    key[i] = k;
}

__device__ __forceinline__ double y_of(const InjectDev& in, double u) {
    const double y = in.height * u;
    return y < in.height ? y : 0.0;   // u < 1, but the product may round up to the height
}

// Re-injection count from the exit counters (main.cpp:235-238); consumes them.
__global__ void k_inject_plan(InjectDev in, const unsigned long long* ab_e, const unsigned long long* ab_i,
                              long long* st) {
    const long long de = (long long)ab_e[0] - st[IS_CONS_E];
    const long long di = (long long)ab_i[0] - st[IS_CONS_I];
    long long n = in.reinject ? de - di : 0;
    if (n < 0) n = 0;
    long long drop = 0;
    if (!(in.x_inj >= 0.0 && in.x_inj < in.width)) { drop = n; n = 0; }   // outside the grid: not inserted
    if (n > in.ecap) { drop += n - in.ecap; n = in.ecap; }
    st[IS_RE] = n;
    st[IS_DROP_NOW] = drop;
    st[IS_CONS_E] = (long long)ab_e[0];
    st[IS_CONS_I] = (long long)ab_i[0];
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_inject_electrons(InjectDev in, GridDev g, const long long* st,
                                                             ParticleBuf se, int32_t* ke) {
    const long long step = st[IS_STEP];
    const int n = (int)st[IS_RE];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        double vx, vy, vz;
        maxwell(in.seed, step, 0, k, 1, in.sig_re, vx, vy, vz);
        put<L>(se, ke, k, in.x_inj, y_of(in, urand(in.seed, step, 0, k, 0)), vx, vy, vz, g);
    }
}

// One workgroup walks the attempts in order, 1024 at a time; an exclusive
// scan of the acceptance flags numbers the accepted ones, and the first
// `target` are kept (electron after the re-injected ones, ion at the same slot).
template <int L>
__global__ __launch_bounds__(kPairBlock) void k_inject_pairs(InjectDev in, GridDev g, long long* st,
                                                             ParticleBuf se, int32_t* ke, ParticleBuf si,
                                                             int32_t* ki) {
    __shared__ int wsum[kPairBlock / 64];
    const long long step = st[IS_STEP];
    const int ne = (int)st[IS_RE];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int base = 0;
    for (int a0 = 0; a0 < in.attempts && base < in.target; a0 += kPairBlock) {
        const int a = a0 + (int)threadIdx.x;
        double x = 0.0;
        bool acc = false;
        if (a < in.attempts) {
            const double u = urand(in.seed, step, 1, a, 0);
            x = in.xm + asin(2.0 * u - 1.0) * (in.x22 - in.x11) / kPi;
            acc = x > 0.0 && x < in.width;
        }
        const unsigned long long m = __ballot(acc);
        const int before = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[wave] = __popcll(m);
        __syncthreads();
        int off = 0, total = 0;
        for (int w = 0; w < kPairBlock / 64; ++w) {
            const int v = wsum[w];
            off += w < wave ? v : 0;
            total += v;
        }
        const int slot = base + off + before;
        if (acc && slot < in.target) {
            const double y = y_of(in, urand(in.seed, step, 1, a, 1));
            double vx, vy, vz;
            maxwell(in.seed, step, 1, a, 2, in.sig_e, vx, vy, vz);
            put<L>(se, ke, ne + slot, x, y, vx, vy, vz, g);
            maxwell(in.seed, step, 1, a, 6, in.sig_i, vx, vy, vz);
            put<L>(si, ki, slot, x, y, vx, vy, vz, g);
        }
        base += total;
        __syncthreads();   // wsum is rewritten by the next round
    }
    if (threadIdx.x == 0) st[IS_NP] = base < in.target ? base : in.target;
}

// Counts per species, clamped to the storage capacity; totals; live counts.
__global__ void k_inject_fin(long long* st, const int32_t* ndev_e, const int32_t* ndev_i, long long cap_e,
                             long long cap_i) {
    const long long re = st[IS_RE], np = st[IS_NP];
    long long n_e = re + np, n_i = np, drop = st[IS_DROP_NOW];
    const long long live_e = *ndev_e, live_i = *ndev_i;
    if (n_e > cap_e - live_e) { drop += n_e - (cap_e - live_e); n_e = cap_e - live_e; }
    if (n_i > cap_i - live_i) { drop += n_i - (cap_i - live_i); n_i = cap_i - live_i; }
    st[IS_NE] = n_e;
    st[IS_NI] = n_i;
    st[IS_LIVE_E] = live_e;
    st[IS_LIVE_I] = live_i;
    st[IS_TOT_RE] += re;
    st[IS_TOT_NP] += np;
    st[IS_TOT_DROP] += drop;
    st[IS_STEP] += 1;
}

// slot of each staged particle inside its cell (after the pushed particles)
__global__ __launch_bounds__(kBlock) void k_inject_count(const long long* st, int which, const int32_t* key,
                                                         int32_t* counts, int32_t* rank) {
    const int n = (int)st[which];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        rank[k] = atomicAdd(&counts[key[k]], 1);
}

template <int L>
__device__ __forceinline__ void copy_particle(const ParticleBuf& src, int k, const ParticleBuf& dst, int d) {
    using P = typename Lay<L>::P;
    using V = typename Lay<L>::V;
    ((P*)dst.pa)[d] = ((const P*)src.pa)[k];
    ((P*)dst.pb)[d] = ((const P*)src.pb)[k];
    ((V*)dst.vx)[d] = ((const V*)src.vx)[k];
    ((V*)dst.vy)[d] = ((const V*)src.vy)[k];
    ((V*)dst.vz)[d] = ((const V*)src.vz)[k];
    dst.cell[d] = src.cell[k];
}

template <int L>
__global__ __launch_bounds__(kBlock) void k_inject_scatter(const long long* st, int which, ParticleBuf src,
                                                           const int32_t* key, const int32_t* rank,
                                                           const int32_t* start, ParticleBuf dst) {
    const int n = (int)st[which];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        copy_particle<L>(src, k, dst, start[key[k]] + rank[k]);
}

// Standalone injection: staged particles appended after the live ones.
template <int L>
__global__ __launch_bounds__(kBlock) void k_inject_append(const long long* st, int which, int live,
                                                          ParticleBuf src, ParticleBuf dst) {
    const int n = (int)st[which];
    const int base = (int)st[live];
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        copy_particle<L>(src, k, dst, base + k);
}

__global__ void k_inject_init(long long* st, const unsigned long long* ab_e, const unsigned long long* ab_i) {
    for (int i = 0; i < IS_N; ++i) st[i] = 0;
    st[IS_CONS_E] = (long long)ab_e[0];
    st[IS_CONS_I] = (long long)ab_i[0];
}

// ------------------------------------------------------------------ host
static constexpr unsigned kStageBlocks = 256;   // grid-stride launches over staged particles

static InjectDev inject_dev(const pic_ctx* c) {
    const Injection& q = c->inj;
    InjectDev d;
    d.x_inj = q.cfg.x_inject;
    d.sig_re = std::sqrt(std::fabs(c->sp[0].qm) * q.cfg.te_reinject_eV);
    d.sig_e = std::sqrt(std::fabs(c->sp[0].qm) * q.cfg.te_eV);
    d.sig_i = std::sqrt(std::fabs(c->sp[1].qm) * q.cfg.ti_eV);
    d.x11 = q.cfg.x11;
    d.x22 = q.cfg.x22;
    d.xm = q.cfg.xm;
    d.width = c->g.width;
    d.height = c->g.height;
    d.reinject = q.cfg.reinject ? 1 : 0;
    d.target = q.target;
    d.attempts = q.attempts;
    d.ecap = q.ecap;
    d.seed = q.cfg.seed;
    return d;
}

static void free_stage(Injection& q) {
    for (int s = 0; s < 2; ++s) {
        ParticleBuf& b = q.stage[s];
        void* ptrs[] = {b.pa, b.pb, b.vx, b.vy, b.vz, b.cell, q.key[s], q.rank[s]};
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        b = ParticleBuf{};
        q.key[s] = q.rank[s] = nullptr;
    }
    if (q.state) (void)hipFree(q.state);
    if (q.host) (void)hipHostFree(q.host);
    if (q.ev) (void)hipEventDestroy(q.ev);
    q.state = nullptr;
    q.host = nullptr;
    q.ev = nullptr;
}

void inject_free(pic_ctx* c) {
    free_stage(c->inj);
    c->inj.on = 0;
    c->inj.pending = 0;
}

// Pairs per step (utils.cpp:619-627), evaluated in the reference's order.
static double pair_rate(const pic_ctx* c, const pic_injection& q) {
    return c->g.height * c->p.dt * (2.0 / kPi) * (q.x22 - q.x11) * q.ionimax * (1.0 / (1.67e6));
}

int inject_setup(pic_ctx* c, const pic_injection* cfg) {
    PIC_TRY(inject_settle(c));
    PIC_HIP(hipStreamSynchronize(c->stream));
    free_stage(c->inj);
    c->inj.on = 0;
    if (!cfg) return PIC_OK;
    if (c->sp.size() < 2) {
        set_error("pic_set_injection: needs species 0 (electrons) and 1 (ions)");
        return PIC_ERR_STATE;
    }
    const pic_injection& q = *cfg;
    const double vals[] = {q.x_inject, q.te_reinject_eV, q.te_eV, q.ti_eV, q.x11, q.x22, q.xm, q.ionimax};
    for (double v : vals)
        if (!std::isfinite(v)) {
            set_error("pic_set_injection: non-finite parameter");
            return PIC_ERR_ARG;
        }
    if (q.te_reinject_eV < 0 || q.te_eV < 0 || q.ti_eV < 0 || q.max_reinject < 0) {
        set_error("pic_set_injection: negative temperature or max_reinject");
        return PIC_ERR_ARG;
    }
    Injection& in = c->inj;
    in.cfg = q;
    const double fn = q.pairs ? pair_rate(c, q) : 0.0;
    if (fn > (double)(1 << 24)) {
        set_error("pic_set_injection: %g pairs per step exceeds 2^24", fn);
        return PIC_ERR_RANGE;
    }
    in.target = fn > 0.0 ? (int)std::ceil(fn) : 0;     // fn <= 0: no pairs (utils.cpp:621-624)
    in.attempts = in.target > 0 ? std::max(10 * in.target, 100) : 0;
    in.ecap = q.reinject ? (int)std::min<int64_t>(q.max_reinject, c->sp[0].capacity) : 0;
    const size_t n0 = ((size_t)in.ecap + (size_t)in.target + 63) & ~(size_t)63;
    const size_t n1 = ((size_t)in.target + 63) & ~(size_t)63;
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    const size_t ps = f64 ? sizeof(double) : sizeof(float);
    bool ok = true;
    for (int s = 0; s < 2; ++s) {
        const size_t n = std::max<size_t>(s == 0 ? n0 : n1, 64);
        ParticleBuf& b = in.stage[s];
        ok = ok && hipMalloc(&b.pa, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.pb, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vx, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vy, n * ps) == hipSuccess;
        ok = ok && hipMalloc(&b.vz, n * ps) == hipSuccess;
        ok = ok && hipMalloc((void**)&b.cell, n * sizeof(int32_t)) == hipSuccess;
        ok = ok && hipMalloc((void**)&in.key[s], n * sizeof(int32_t)) == hipSuccess;
        ok = ok && hipMalloc((void**)&in.rank[s], n * sizeof(int32_t)) == hipSuccess;
    }
    ok = ok && hipMalloc((void**)&in.state, IS_N * sizeof(long long)) == hipSuccess;
    ok = ok && hipHostMalloc((void**)&in.host, IS_N * sizeof(long long), hipHostMallocDefault) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&in.ev, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        free_stage(in);
        set_error("pic_set_injection: allocation of the staging buffers failed");
        return PIC_ERR_ALLOC;
    }
    hipLaunchKernelGGL(k_inject_init, dim3(1), dim3(1), 0, c->stream, in.state, c->sp[0].absorbed,
                       c->sp[1].absorbed);
    PIC_HIP(hipGetLastError());
    PIC_HIP(hipStreamSynchronize(c->stream));
    in.on = 1;
    return PIC_OK;
}

int inject_settle(pic_ctx* c) {
    Injection& q = c->inj;
    if (!q.pending) return PIC_OK;
    q.pending = 0;
    PIC_HIP(hipEventSynchronize(q.ev));
    c->sp[0].n_upper = q.host[IS_LIVE_E] + q.host[IS_NE];
    c->sp[1].n_upper = q.host[IS_LIVE_I] + q.host[IS_NI];
    return PIC_OK;
}

// This step's new particles into the staging buffers.  Inside pic_step it
// runs after the push of both species (their exits are counted) and before
// their scans; the live counts it clamps against are those before the push,
// an upper bound of the pushed ones.
int inject_generate(pic_ctx* c) {
    Injection& q = c->inj;
    if (!q.on) {
        set_error("injection is not configured (pic_set_injection)");
        return PIC_ERR_STATE;
    }
    const InjectDev d = inject_dev(c);
    Species& e = c->sp[0];
    Species& i = c->sp[1];
    hipLaunchKernelGGL(k_inject_plan, dim3(1), dim3(1), 0, c->stream, d, e.absorbed, i.absorbed, q.state);
    const bool f64 = c->p.layout == PIC_LAYOUT_F64;
    if (q.ecap > 0) {
        const unsigned nb = std::min<unsigned>(blocks_for(q.ecap), kStageBlocks);
        if (f64)
            hipLaunchKernelGGL(k_inject_electrons<PIC_LAYOUT_F64>, dim3(nb), dim3(kBlock), 0, c->stream, d, c->g,
                               q.state, q.stage[0], q.key[0]);
        else
            hipLaunchKernelGGL(k_inject_electrons<PIC_LAYOUT_CELL32>, dim3(nb), dim3(kBlock), 0, c->stream, d,
                               c->g, q.state, q.stage[0], q.key[0]);
    }
    if (f64)
        hipLaunchKernelGGL(k_inject_pairs<PIC_LAYOUT_F64>, dim3(1), dim3(kPairBlock), 0, c->stream, d, c->g,
                           q.state, q.stage[0], q.key[0], q.stage[1], q.key[1]);
    else
        hipLaunchKernelGGL(k_inject_pairs<PIC_LAYOUT_CELL32>, dim3(1), dim3(kPairBlock), 0, c->stream, d, c->g,
                           q.state, q.stage[0], q.key[0], q.stage[1], q.key[1]);
    hipLaunchKernelGGL(k_inject_fin, dim3(1), dim3(1), 0, c->stream, q.state, e.ndev + e.cur, i.ndev + i.cur,
                       (long long)e.capacity, (long long)i.capacity);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

static unsigned stage_blocks(const Injection& q, int s) {
    const int64_t n = s == 0 ? (int64_t)q.ecap + q.target : (int64_t)q.target;
    return std::max(1u, std::min<unsigned>(blocks_for(n), kStageBlocks));
}

int inject_count(pic_ctx* c, int s) {
    Injection& q = c->inj;
    hipLaunchKernelGGL(k_inject_count, dim3(stage_blocks(q, s)), dim3(kBlock), 0, c->stream, q.state,
                       s == 0 ? (int)IS_NE : (int)IS_NI, q.key[s], c->sp[s].counts, q.rank[s]);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

int inject_scatter(pic_ctx* c, int s, int dst) {
    Injection& q = c->inj;
    Species& sp = c->sp[s];
    const int which = s == 0 ? (int)IS_NE : (int)IS_NI;
    if (c->p.layout == PIC_LAYOUT_F64)
        hipLaunchKernelGGL(k_inject_scatter<PIC_LAYOUT_F64>, dim3(stage_blocks(q, s)), dim3(kBlock), 0, c->stream,
                           q.state, which, q.stage[s], q.key[s], q.rank[s], sp.start, sp.buf[dst]);
    else
        hipLaunchKernelGGL(k_inject_scatter<PIC_LAYOUT_CELL32>, dim3(stage_blocks(q, s)), dim3(kBlock), 0,
                           c->stream, q.state, which, q.stage[s], q.key[s], q.rank[s], sp.start, sp.buf[dst]);
    PIC_HIP(hipGetLastError());
    return PIC_OK;
}

// Counters to pinned host memory; the next launch that sizes a grid by the
// live count waits for them (inject_settle).
int inject_publish(pic_ctx* c) {
    Injection& q = c->inj;
    PIC_HIP(hipMemcpyAsync(q.host, q.state, IS_N * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipEventRecord(q.ev, c->stream));
    q.pending = 1;
    return PIC_OK;
}

int launch_key_from_cell(pic_ctx* c, int s, int64_t n);   // pic_particles.hip

// Injection into binned storage outside pic_step: generate, append after the
// live particles, re-bin both species.
int inject_append_rebin(pic_ctx* c) {
    PIC_TRY(inject_settle(c));
    for (int s = 0; s < 2; ++s)
        if (!c->sp[s].sorted) PIC_TRY(launch_bin(c, s));
    PIC_TRY(inject_generate(c));
    Injection& q = c->inj;
    for (int s = 0; s < 2; ++s) {
        Species& sp = c->sp[s];
        const int which = s == 0 ? (int)IS_NE : (int)IS_NI, live = s == 0 ? (int)IS_LIVE_E : (int)IS_LIVE_I;
        if (c->p.layout == PIC_LAYOUT_F64)
            hipLaunchKernelGGL(k_inject_append<PIC_LAYOUT_F64>, dim3(stage_blocks(q, s)), dim3(kBlock), 0,
                               c->stream, q.state, which, live, q.stage[s], sp.buf[sp.cur]);
        else
            hipLaunchKernelGGL(k_inject_append<PIC_LAYOUT_CELL32>, dim3(stage_blocks(q, s)), dim3(kBlock), 0,
                               c->stream, q.state, which, live, q.stage[s], sp.buf[sp.cur]);
        PIC_HIP(hipGetLastError());
    }
    PIC_HIP(hipMemcpyAsync(q.host, q.state, IS_N * sizeof(long long), hipMemcpyDeviceToHost, c->stream));
    PIC_HIP(hipStreamSynchronize(c->stream));
    const int64_t live[2] = {q.host[IS_LIVE_E], q.host[IS_LIVE_I]};
    const int64_t added[2] = {q.host[IS_NE], q.host[IS_NI]};
    for (int s = 0; s < 2; ++s) {
        if (added[s] == 0) continue;   // nothing appended: still binned
        PIC_TRY(launch_key_from_cell(c, s, live[s] + added[s]));
        c->sp[s].deposit_pending = 0;
        PIC_TRY(launch_bin(c, s));
    }
    return PIC_OK;
}

int inject_stats(pic_ctx* c, int64_t* out) {
    Injection& q = c->inj;
    if (!q.state) {
        for (int i = 0; i < 6; ++i) out[i] = 0;
        return PIC_OK;
    }
    long long st[IS_N];
    PIC_HIP(hipStreamSynchronize(c->stream));
    PIC_HIP(hipMemcpy(st, q.state, sizeof(st), hipMemcpyDeviceToHost));
    out[0] = st[IS_STEP];
    out[1] = st[IS_TOT_RE];
    out[2] = st[IS_TOT_NP];
    out[3] = st[IS_TOT_DROP];
    out[4] = st[IS_STEP] ? st[IS_NE] : 0;
    out[5] = st[IS_STEP] ? st[IS_NI] : 0;
    return PIC_OK;
}

} // namespace pic
