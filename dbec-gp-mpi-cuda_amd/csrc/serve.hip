// serve.hip - the continuous-batching request loop of include/dbec_serve.h.
//
// A bucket holds the solves of one (grid, dx, dt, cutoff, par, mode,
// precision): one context with the shared tables, and `cap` slots whose psi,
// spectrum, row partials and scalars sit in slabs (slot s at base + s *
// stride).  A chunk advances every active slot of a bucket by the same number
// of steps with the batched step kernels, then measures the observables of
// each slot through a per-slot view context and copies all slot scalars to
// pinned host memory behind one event.  The host polls the events; a bucket
// whose chunk has finished retires converged solves, admits queued requests
// into free slots and launches its next chunk.  Buckets run on their own
// streams (default priority), so chunks of different sizes overlap on the
// device.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "dbec_serve.h"
#include "core.hpp"
#include "ops.hpp"

namespace dbec {
namespace {

// Gaussian initial state of initpsi (imag3d.cu:857-864): exp(-(g x^2 + n y^2 +
// l z^2) / 2) / cpsi, real or with zero imaginary part.
template <typename R>
__device__ __forceinline__ void put_psi(R *psi, long n, double v, bool cplx) {
    if (cplx) {
        psi[2 * n] = (R)v;
        psi[2 * n + 1] = (R)0;
    } else {
        psi[n] = (R)v;
    }
}

__global__ void k_widen(double *__restrict__ dst, const float *__restrict__ src, long n) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (double)src[t];
}

template <typename R>
__global__ void k_gauss(R *__restrict__ psi, const double *__restrict__ sq, int Nx, int Ny, int Nz, double vg,
                        double vn, double vl, double cpsi, int cplx) {
    const long n = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= (long)Nx * Ny * Nz) return;
    const long row = n / Nz;
    const int k = (int)(n % Nz), i = (int)(row / Ny), j = (int)(row % Ny);
    const double t = exp(-0.5 * (vg * sq[i] + vn * sq[Nx + j] + vl * sq[Nx + Ny + k]));
    put_psi(psi, n, t / cpsi, cplx != 0);
}

double now_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

size_t round_up(size_t b) { return (b + 255) & ~size_t(255); }

// Chunks of a real-time bucket launched back to back before the host looks
// again (admission into the bucket waits at most this many chunks).
constexpr long kChainMax = 8;
// |dmu/mu| below which a mixed ground state whose mu changes stop falling is
// at FP32's rounding floor (mu of an FP32 wave function carries about 1e-7).
constexpr double kPolishFloor = 1e-6;

struct Job {
    dbec_request r;
    double t_submit = 0., t_start = 0.;
    long steps = 0, steps_fp64 = 0;
    double mu_prev = 0.;
    double dmu_prev = -1.;  // |dmu/mu| of the previous check (< 0: none yet)
    bool have_prev = false;
    bool polishing = false;  // moved to an FP64 slot
    bool to_polish = false;  // mixed phase done, waiting for an FP64 slot
};

struct Bucket {
    dbec_request key;  // grid, dx, dt, cutoff, par, mode, prec
    dbec_ctx *tab = nullptr;
    int cap = 0;
    char *psi = nullptr, *spec = nullptr, *rows = nullptr, *scal = nullptr;
    int *d_slot = nullptr;
    SolveParams *d_par = nullptr;
    long psi_st = 0, spec_st = 0, rows_st = 0, scal_st = 0;
    std::vector<dbec_ctx *> views;
    std::vector<std::unique_ptr<Job>> job;  // per slot, empty = free
    int *h_slot = nullptr;
    SolveParams *h_par = nullptr;
    double *h_scal = nullptr;
    hipEvent_t ev = nullptr;
    // recorded on this bucket's stream after a widen that read another
    // bucket's slot (migrate); that bucket's stream waits on it before the
    // slot can be written again
    hipEvent_t ev_mig = nullptr;
    bool inflight = false;
    // chunk graphs: check_every steps of nb solves, by nb (the slot list and
    // parameters are device data the graph reads, so any composition of nb
    // solves replays the same graph)
    std::vector<hipGraphExec_t> graphs;
    std::vector<int> launched;
    double bytes = 0.;
    double last_used = 0.;

    int active() const {
        int a = 0;
        for (auto &j : job) a += j ? 1 : 0;
        return a;
    }
    int free_slot() const {
        for (int i = 0; i < cap; i++)
            if (!job[i]) return i;
        return -1;
    }
};

bool same_bucket(const dbec_request &a, const dbec_request &b) {
    return a.mode == b.mode && a.prec == b.prec && a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.dx == b.dx &&
           a.dy == b.dy && a.dz == b.dz && a.dt == b.dt && a.cutoff == b.cutoff && a.par == b.par;
}

dbec_params params_of(const dbec_request &r, int device) {
    dbec_params p;
    std::memset(&p, 0, sizeof p);
    p.nx = r.nx;
    p.ny = r.ny;
    p.nz = r.nz;
    p.dx = r.dx;
    p.dy = r.dy;
    p.dz = r.dz;
    p.dt = r.dt;
    p.vgamma = r.vgamma;
    p.vnu = r.vnu;
    p.vlambda = r.vlambda;
    p.cutoff = r.cutoff;
    p.par = r.par;
    p.pi = 0.;
    p.device = device;
    return p;
}

}  // namespace
}  // namespace dbec

using namespace dbec;

struct dbec_server {
    dbec_server_config cfg;
    int device = 0;
    std::chrono::steady_clock::time_point t0;
    std::deque<std::unique_ptr<Job>> queue;
    std::vector<std::unique_ptr<Bucket>> buckets;
    std::vector<dbec_result> done;
    long pending = 0;
    double budget = 0.;
    std::string err;

    int fail(const std::string &m) {
        err = m;
        return DBEC_EDEVICE;
    }

    // Device bytes of a bucket with `cap` slots for requests like r.
    static double bucket_bytes(const dbec_request &r, int cap) {
        const double N = (double)r.nx * r.ny * r.nz, Nkz = (double)(r.nz / 2 + 1);
        const double re = r.prec == DBEC_MIXED ? 4. : 8., pe = r.mode == DBEC_REAL ? 2 * re : re;
        const double slot = round_up((size_t)(N * pe)) + round_up((size_t)(Nkz * r.nx * r.ny * 2 * re)) +
                            round_up((size_t)(2. * r.nx * r.ny * 8.)) + SC_COUNT * 8.;
        const double tables = N * pe + Nkz * r.nx * r.ny * 2 * re + N * re + 2. * r.nx * r.ny * 8. +
                              Nkz * (r.nx / 2 + 1) * (r.ny / 2 + 1) * re + 3. * 16. * (r.nx + r.ny + r.nz) * 3;
        return tables + cap * slot;
    }

    double held() const {
        double b = 0.;
        for (auto &k : buckets) b += k->bytes;
        return b;
    }

    void release(Bucket *b) {
        (void)hipStreamSynchronize(b->tab->stream);
        for (auto *v : b->views) {
            v->view = true;
            (void)dbec_destroy(v);
        }
        (void)hipFree(b->psi);
        (void)hipFree(b->spec);
        (void)hipFree(b->rows);
        (void)hipFree(b->scal);
        (void)hipFree(b->d_slot);
        (void)hipFree(b->d_par);
        (void)hipHostFree(b->h_slot);
        (void)hipHostFree(b->h_par);
        (void)hipHostFree(b->h_scal);
        if (b->ev) (void)hipEventDestroy(b->ev);
        if (b->ev_mig) (void)hipEventDestroy(b->ev_mig);
        for (auto g : b->graphs)
            if (g) (void)hipGraphExecDestroy(g);
        (void)dbec_destroy(b->tab);
    }

    int make_bucket(const dbec_request &r, Bucket **out) {
        auto b = std::make_unique<Bucket>();
        b->key = r;
        b->cap = cfg.max_batch;
        const dbec_params p = params_of(r, device);
        int rc = dbec_create(&p, (dbec_mode)r.mode, (dbec_prec)r.prec, nullptr, &b->tab);
        if (rc != DBEC_OK) {
            err = dbec_last_error(nullptr);
            return rc;
        }
        dbec_ctx *t = b->tab;
        b->psi_st = (long)round_up((size_t)t->N * core_psi_elem(t));
        b->spec_st = (long)round_up((size_t)t->Nkz * t->Nx * t->Ny * 2 * core_real_elem(t));
        b->rows_st = (long)round_up((size_t)2 * t->Nx * t->Ny * sizeof(double));
        b->scal_st = SC_COUNT * sizeof(double);
        const int cap = b->cap;
        hipError_t e = hipSuccess;
        auto dm = [&](void **ptr, size_t bytes) {
            if (e == hipSuccess) e = hipMalloc(ptr, bytes);
        };
        dm((void **)&b->psi, (size_t)cap * b->psi_st);
        dm((void **)&b->spec, (size_t)cap * b->spec_st);
        dm((void **)&b->rows, (size_t)cap * b->rows_st);
        dm((void **)&b->scal, (size_t)cap * b->scal_st);
        dm((void **)&b->d_slot, (size_t)cap * sizeof(int));
        dm((void **)&b->d_par, (size_t)cap * sizeof(SolveParams));
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_slot, cap * sizeof(int), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_par, cap * sizeof(SolveParams), 0);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_scal, (size_t)cap * b->scal_st, 0);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_mig, hipEventDisableTiming);
        if (e != hipSuccess) {
            release(b.get());
            return fail(std::string("bucket allocation: ") + hipGetErrorString(e));
        }
        std::memset(b->h_par, 0, cap * sizeof(SolveParams));
        b->job.resize(cap);
        b->views.resize(cap, nullptr);
        for (int s = 0; s < cap; s++) {
            dbec_ctx *v = new dbec_ctx();
            v->p = t->p;
            v->mode = t->mode;
            v->prec = t->prec;
            v->device = t->device;
            v->stream = t->stream;
            v->own_stream = false;
            v->Nx = t->Nx;
            v->Ny = t->Ny;
            v->Nz = t->Nz;
            v->N = t->N;
            v->Nkz = t->Nkz;
            v->pi = t->pi;
            std::memcpy(v->A0r, t->A0r, sizeof v->A0r);
            std::memcpy(v->A, t->A, sizeof v->A);
            std::memcpy(v->mA, t->mA, sizeof v->mA);
            v->tune = t->tune;
            v->b = t->b;  // tables, phi and reduction scratch of the bucket
            v->b.psi = b->psi + (size_t)s * b->psi_st;
            v->b.spec = b->spec + (size_t)s * b->spec_st;
            v->b.rows = reinterpret_cast<double *>(b->rows + (size_t)s * b->rows_st);
            v->b.scal = reinterpret_cast<double *>(b->scal + (size_t)s * b->scal_st);
            v->view = true;
            b->views[s] = v;
        }
        b->bytes = bucket_bytes(r, cap);
        *out = b.get();
        buckets.push_back(std::move(b));
        return DBEC_OK;
    }

    // Room for a new bucket: release idle buckets, least recently used first.
    bool make_room(double need) {
        while (held() + need > budget) {
            Bucket *victim = nullptr;
            for (auto &b : buckets)
                if (!b->inflight && b->active() == 0 && (!victim || b->last_used < victim->last_used)) victim = b.get();
            if (!victim) return false;
            release(victim);
            buckets.erase(std::find_if(buckets.begin(), buckets.end(),
                                       [&](const std::unique_ptr<Bucket> &x) { return x.get() == victim; }));
        }
        return true;
    }

    // trap and couplings of the solve in slot s
    void set_params(Bucket *b, int s, const dbec_request &r) {
        dbec_ctx *v = b->views[s];
        v->p.vgamma = r.vgamma;
        v->p.vnu = r.vnu;
        v->p.vlambda = r.vlambda;
        v->g = r.g;
        v->gd = r.gd;
        b->h_par[s].tc = trap_const(v);
        b->h_par[s].g = r.g;
        b->h_par[s].gd = r.gd;
    }

    Bucket *find_bucket(const dbec_request &r) {
        for (auto &k : buckets)
            if (same_bucket(k->key, r)) return k.get();
        return nullptr;
    }

    // A mixed ground state that reached polish_tol continues in FP64: its
    // (renormalised) psi is widened into a free slot of the FP64 bucket of
    // the same grid, on that bucket's stream; the mixed bucket's chunk has
    // completed, so the source is final.
    int migrate(Bucket *from, int s) {
        dbec_request key = from->job[s]->r;
        key.prec = DBEC_FP64;
        Bucket *b = find_bucket(key);
        if (!b) {
            if (!make_room(bucket_bytes(key, cfg.max_batch))) return DBEC_OK;  // retry later
            int rc = make_bucket(key, &b);
            if (rc != DBEC_OK) return rc;
        }
        const int d = b->inflight ? -1 : b->free_slot();
        if (d < 0) return DBEC_OK;
        dbec_ctx *v = b->views[d];
        set_params(b, d, key);
        const long n = v->N;
        k_widen<<<(unsigned)((n + 255) / 256), 256, 0, v->stream>>>(static_cast<double *>(v->b.psi),
                                                                     static_cast<const float *>(from->views[s]->b.psi), n);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_SHIFT, 0.0);
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_CAPPED, 0.0);
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_SCALE, 1.0);
        // slot s of the source bucket is free once this returns: a request
        // admitted into it writes it on the source bucket's stream, which
        // must first wait for the widen to have read it
        if (e == hipSuccess) e = hipEventRecord(b->ev_mig, v->stream);
        if (e == hipSuccess) e = hipStreamWaitEvent(from->tab->stream, b->ev_mig, 0);
        if (e != hipSuccess) return fail(std::string("polish: ") + hipGetErrorString(e));
        std::unique_ptr<Job> j = std::move(from->job[s]);
        j->polishing = true;
        j->to_polish = false;
        j->have_prev = false;
        b->job[d] = std::move(j);
        return DBEC_OK;
    }

    int admit(Bucket *b, int s, std::unique_ptr<Job> j) {
        dbec_ctx *v = b->views[s];
        const dbec_request &r = j->r;
        set_params(b, s, r);
        const double pi = 3.14159265358979;
        const double cpsi = std::sqrt(pi * std::sqrt(pi / (r.vgamma * r.vnu * r.vlambda)));
        const dim3 blk(256), grd((unsigned)((v->N + 255) / 256));
        const int cplx = v->mode == DBEC_REAL;
        if (v->prec == DBEC_MIXED)
            k_gauss<float><<<grd, blk, 0, v->stream>>>(static_cast<float *>(v->b.psi), v->b.sq, (int)v->Nx, (int)v->Ny,
                                                        (int)v->Nz, r.vgamma, r.vnu, r.vlambda, cpsi, cplx);
        else
            k_gauss<double><<<grd, blk, 0, v->stream>>>(static_cast<double *>(v->b.psi), v->b.sq, (int)v->Nx,
                                                         (int)v->Ny, (int)v->Nz, r.vgamma, r.vnu, r.vlambda, cpsi, cplx);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_SHIFT, 0.0);
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_CAPPED, 0.0);
        if (e == hipSuccess) e = launch_set_scalar(v->stream, v->b.scal + SC_SCALE, 1.0);
        if (e == hipSuccess && v->mode == DBEC_IMAG) {
            // imag3d normalises psi0 (imag3d.cu:239): the first step reads psi / norm
            double *r0 = v->b.rows;
            e = launch_rows_mod2(v->stream, v, r0, nullptr);
            if (e == hipSuccess) e = launch_finalize(v->stream, v, 0, r0, nullptr);
        }
        if (e != hipSuccess) return fail(std::string("admission: ") + hipGetErrorString(e));
        j->t_start = now_since(t0);
        b->job[s] = std::move(j);
        return DBEC_OK;
    }

    int launch_chunk(Bucket *b) {
        dbec_ctx *t = b->tab;
        const bool imag = t->mode == DBEC_IMAG;
        b->launched.clear();
        long K = cfg.check_every;
        for (int s = 0; s < b->cap; s++) {
            if (!b->job[s] || b->job[s]->to_polish) continue;
            b->launched.push_back(s);
            if (!imag) K = std::min(K, b->job[s]->r.max_steps - b->job[s]->steps);
            else K = std::min(K, std::max(1L, b->job[s]->r.max_steps - b->job[s]->steps));
        }
        const int na = (int)b->launched.size();
        if (na == 0) return DBEC_OK;
        // Real-time solves need no decision between chunks: while every
        // active one has at least R more chunks to run, R chunks go out
        // back to back (at most kChainMax), so the device does not wait for
        // the host between them.  Ground states are checked every chunk.
        int R = 1;
        if (!imag && K == cfg.check_every) {
            long rmin = kChainMax;
            for (int s : b->launched) rmin = std::min(rmin, (b->job[s]->r.max_steps - b->job[s]->steps) / K);
            R = (int)std::max(1L, rmin);
        }
        const bool fp64 = t->prec == DBEC_FP64;
        for (int i = 0; i < na; i++) b->h_slot[i] = b->launched[i];
        hipStream_t st = t->stream;
        hipError_t e = hipMemcpyAsync(b->d_slot, b->h_slot, na * sizeof(int), hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->d_par, b->h_par, b->cap * sizeof(SolveParams), hipMemcpyHostToDevice, st);
        if (e != hipSuccess) return fail(std::string("chunk upload: ") + hipGetErrorString(e));
        Solves v;
        v.psi = b->psi;
        v.spec = b->spec;
        v.rows = reinterpret_cast<double *>(b->rows);
        v.scal = reinterpret_cast<double *>(b->scal);
        v.psi_st = b->psi_st;
        v.spec_st = b->spec_st;
        v.rows_st = b->rows_st;
        v.scal_st = b->scal_st;
        v.slot = b->d_slot;
        v.par = b->d_par;
        v.p0 = b->h_par[b->launched[0]];
        v.nb = na;
        v.use_scale = imag ? 1 : 0;
        v.shift = (imag && t->prec == DBEC_MIXED) ? 1 : 0;
        v.fold = 0;
        // steps after the first fold the previous norm in; with the z sweep
        // fused into the next forward z pass, the y pass forms that norm
        const bool fuse = core_fuse_z(t);
        Solves vf = v;
        vf.fold = imag ? (fuse ? 2 : 1) : 0;
        static const bool use_graphs = [] {
            const char *e = getenv("DBEC_SERVE_GRAPHS");
            return !e || atoi(e) != 0;
        }();
        if (use_graphs && K == cfg.check_every) {
            if ((int)b->graphs.size() <= na) b->graphs.resize(na + 1, nullptr);
            if (!b->graphs[na]) {
                // the step kernels only launch: no allocation, copy or memset
                hipGraph_t g = nullptr;
                e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
                int rc = DBEC_OK;
                for (long k = 0; k < K && e == hipSuccess && rc == DBEC_OK; k++)
                    rc = core_step(t, k ? vf : v, fuse && k > 0, fuse && k + 1 < K);
                if (imag && rc == DBEC_OK) rc = core_norm(t, v);
                hipError_t e2 = hipStreamEndCapture(st, &g);
                if (e == hipSuccess) e = e2;
                if (rc == DBEC_OK && e == hipSuccess) e = hipGraphInstantiate(&b->graphs[na], g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);  // on every path: the exec graph keeps what it needs
                if (rc != DBEC_OK) return fail(std::string("step: ") + t->err);
                if (e != hipSuccess) {
                    b->graphs[na] = nullptr;
                    return fail(std::string("chunk graph: ") + hipGetErrorString(e));
                }
            }
            for (int c = 0; c < R && e == hipSuccess; c++) e = hipGraphLaunch(b->graphs[na], st);
            if (e != hipSuccess) return fail(std::string("chunk graph launch: ") + hipGetErrorString(e));
        } else {
            for (int c = 0; c < R; c++)
                for (long k = 0; k < K; k++) {
                    int rc = core_step(t, k ? vf : v, fuse && k > 0, fuse && k + 1 < K);
                    if (rc != DBEC_OK) return fail(std::string("step: ") + t->err);
                }
            if (imag) {
                int rc = core_norm(t, v);
                if (rc != DBEC_OK) return fail(std::string("norm: ") + t->err);
            }
        }
        for (int s : b->launched) {
            Job &j = *b->job[s];
            j.steps += K * R;
            if (fp64 && j.polishing) j.steps_fp64 += K * R;
            int what = 0;
            if (imag) what = OBS_MUEN | OBS_RMS;
            else if (j.steps >= j.r.max_steps) what = OBS_MUEN | OBS_RMS;
            if (what) {
                int rc = core_observe_async(b->views[s], what);
                if (rc != DBEC_OK) return fail(std::string("observables: ") + b->views[s]->err);
            }
        }
        e = hipMemcpyAsync(b->h_scal, b->scal, (size_t)b->cap * b->scal_st, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(b->ev, st);
        if (e != hipSuccess) return fail(std::string("chunk end: ") + hipGetErrorString(e));
        b->inflight = true;
        b->last_used = now_since(t0);
        return DBEC_OK;
    }

    void retire(Bucket *b) {
        const bool imag = b->tab->mode == DBEC_IMAG;
        const double t = now_since(t0);
        for (int s : b->launched) {
            Job &j = *b->job[s];
            const double *sc = b->h_scal + (size_t)s * SC_COUNT;
            int status = -1;
            if (imag) {
                const double mu = sc[SC_MU];
                const double dmu = std::fabs(mu - j.mu_prev);
                const bool polish = b->tab->prec == DBEC_MIXED && cfg.polish_tol > j.r.mutol;
                const double rel = dmu / std::fabs(mu);
                // FP32 resolves mu to about 1e-7: once |dmu/mu| is below
                // kPolishFloor and no longer falls between checks, the mixed
                // phase is at its rounding floor and moves on even if
                // polish_tol is below that floor
                const bool floor = j.dmu_prev >= 0. && rel < kPolishFloor && rel >= j.dmu_prev;
                if (!std::isfinite(mu) || sc[SC_CAPPED] != 0.) status = DBEC_SOLVE_DIVERGED;
                else if (polish && j.have_prev && (rel <= cfg.polish_tol || floor)) j.to_polish = true;
                else if (j.have_prev && dmu <= j.r.mutol * std::fabs(mu)) status = DBEC_SOLVE_CONVERGED;
                else if (j.steps >= j.r.max_steps) status = DBEC_SOLVE_NOT_CONVERGED;
                j.dmu_prev = j.have_prev ? rel : -1.;
                j.mu_prev = mu;
                j.have_prev = true;
            } else if (j.steps >= j.r.max_steps) {
                status = std::isfinite(sc[SC_MU]) ? DBEC_SOLVE_STEPS_DONE : DBEC_SOLVE_DIVERGED;
            }
            if (status < 0) continue;
            dbec_result res;
            std::memset(&res, 0, sizeof res);
            res.id = j.r.id;
            res.status = status;
            res.steps = j.steps;
            res.steps_fp64 = j.steps_fp64;
            res.norm = imag ? sc[SC_LASTNORM] : sc[SC_NORM];
            res.mu = sc[SC_MU];
            res.en = sc[SC_EN];
            for (int q = 0; q < 4; q++) res.rms[q] = sc[SC_R0 + q];
            res.t_submit = j.t_submit;
            res.t_start = j.t_start;
            res.t_done = t;
            done.push_back(res);
            b->job[s].reset();
            pending--;
        }
        b->inflight = false;
    }

    int progress(int wait, dbec_result *out, int max, int *n) {
        (void)hipSetDevice(device);
        bool changed = false;
        for (auto &b : buckets) {
            if (!b->inflight) continue;
            hipError_t q = hipEventQuery(b->ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return fail(std::string("chunk: ") + hipGetErrorString(q));
            retire(b.get());
            changed = true;
        }
        // mixed ground states that wait for an FP64 slot
        for (size_t bi = 0; bi < buckets.size(); bi++) {
            Bucket *b = buckets[bi].get();
            if (b->inflight) continue;
            for (int s = 0; s < b->cap; s++) {
                if (!b->job[s] || !b->job[s]->to_polish) continue;
                int rc = migrate(b, s);
                if (rc != DBEC_OK) return rc;
                if (!b->job[s]) changed = true;
            }
        }
        // admission, in submission order
        for (auto it = queue.begin(); it != queue.end();) {
            const dbec_request &r = (*it)->r;
            Bucket *b = nullptr;
            for (auto &k : buckets)
                if (same_bucket(k->key, r)) b = k.get();
            if (!b) {
                const double need = bucket_bytes(r, cfg.max_batch);
                if (!make_room(need)) {
                    ++it;
                    continue;
                }
                int rc = make_bucket(r, &b);
                if (rc != DBEC_OK) return rc;
            }
            int s = b->inflight ? -1 : b->free_slot();
            if (s < 0) {
                ++it;
                continue;
            }
            int rc = admit(b, s, std::move(*it));
            if (rc != DBEC_OK) return rc;
            it = queue.erase(it);
            changed = true;
        }
        for (auto &b : buckets) {
            if (b->inflight || b->active() == 0) continue;
            int rc = launch_chunk(b.get());
            if (rc != DBEC_OK) return rc;
            changed = true;
        }
        int m = 0;
        while (m < max && !done.empty()) {
            out[m++] = done.front();
            done.erase(done.begin());
        }
        *n = m;
        if (wait && !changed && m == 0) std::this_thread::sleep_for(std::chrono::microseconds(200));
        return DBEC_OK;
    }
};

extern "C" {

int dbec_server_create(const dbec_server_config *cfg, dbec_server **out) {
    if (!cfg || !out) return DBEC_EINVAL;
    *out = nullptr;
    if (cfg->max_batch < 1 || cfg->max_batch > 64 || cfg->check_every < 1) return DBEC_EINVAL;
    auto s = std::make_unique<dbec_server>();
    s->cfg = *cfg;
    int dev = cfg->device;
    if (dev < 0) {
        if (hipGetDevice(&dev) != hipSuccess) return DBEC_EDEVICE;
    }
    if (hipSetDevice(dev) != hipSuccess) return DBEC_EDEVICE;
    s->device = dev;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return DBEC_EDEVICE;
    s->budget = cfg->mem_budget > 0 ? cfg->mem_budget : 0.5 * (double)fr;
    s->t0 = std::chrono::steady_clock::now();
    *out = s.release();
    return DBEC_OK;
}

int dbec_server_destroy(dbec_server *s) {
    if (!s) return DBEC_OK;
    (void)hipSetDevice(s->device);
    for (auto &b : s->buckets) s->release(b.get());
    delete s;
    return DBEC_OK;
}

double dbec_server_time(const dbec_server *s) { return s ? now_since(s->t0) : 0.; }

int dbec_server_submit(dbec_server *s, const dbec_request *r) {
    if (!s || !r) return DBEC_EINVAL;
    if (r->mode != DBEC_IMAG && r->mode != DBEC_REAL) return s->fail("bad mode"), DBEC_EINVAL;
    if (r->prec != DBEC_FP64 && r->prec != DBEC_MIXED) return s->fail("bad precision"), DBEC_EINVAL;
    if (r->max_steps < 1 || !(r->dt > 0) || (r->mode == DBEC_IMAG && !(r->mutol > 0)))
        return s->fail("bad step count, dt or mutol"), DBEC_EINVAL;
    if (!fft_len_supported(r->nx) || !fft_len_supported(r->ny) || !fft_len_supported(r->nz))
        return s->fail("unsupported grid size"), DBEC_EUNSUPPORTED;
    if (dbec_server::bucket_bytes(*r, s->cfg.max_batch) > s->budget)
        return s->fail("request does not fit the memory budget"), DBEC_ENOMEM;
    auto j = std::make_unique<Job>();
    j->r = *r;
    j->t_submit = now_since(s->t0);
    s->queue.push_back(std::move(j));
    s->pending++;
    return DBEC_OK;
}

int dbec_server_progress(dbec_server *s, int wait, dbec_result *out, int max, int *n) {
    if (!s || !n || (max > 0 && !out)) return DBEC_EINVAL;
    *n = 0;
    return s->progress(wait, out, max, n);
}

long dbec_server_pending(const dbec_server *s) { return s ? s->pending : 0; }

int dbec_server_usage(const dbec_server *s, double *bytes, int *buckets) {
    if (!s) return DBEC_EINVAL;
    if (bytes) *bytes = s->held();
    if (buckets) *buckets = (int)s->buckets.size();
    return DBEC_OK;
}

const char *dbec_server_last_error(const dbec_server *s) { return s ? s->err.c_str() : "null server"; }

}  // extern "C"
