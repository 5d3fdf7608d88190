/*
 * organoid.hip -- host side of the C-ABI declared in include/organoid.h.
 *
 * One org_ctx owns the device state of one organoid and one HIP stream
 * (default priority).  FP32 contexts with the 128-point FFT box
 * (zone_size 65) run the kernels of organoid_kernels.hip; FP64 contexts and
 * other box sizes run the per-function full-plane path of
 * organoid_generic.hip (ctx->gen), behind the same entry points.
 */
#include <mutex>

#include "organoid_kernels.hip"
#include "organoid_fno.h"
#include "organoid_generic.h"

namespace {

std::atomic<int> g_last_error{0};

int note_error(int code) {
    int zero = 0;
    if (code) g_last_error.compare_exchange_strong(zero, code);
    return code;
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            fprintf(stderr, "organoid: %s failed: %s (%s:%d)\n", #expr, hipGetErrorString(_e), \
                    __FILE__, __LINE__);                                               \
            return note_error(ORG_EHIP);                                               \
        }                                                                              \
    } while (0)

#define TRY(x) do { int _r = (x); if (_r) return _r; } while (0)

/* steps per replay of the batched step's graph (the launch of a graph costs
 * more than the gaps between its kernels) */
#define ORG_GRAPH_STEPS 8
/* largest workgroup count of a non-cell role of k_pre / k_mid / k_post */
#define ORG_ROLE_MAX 512

}  // namespace

struct org_ctx {
    org_params p{};
    DevP dp{};
    hipStream_t stream = nullptr;
    gen::Generic *gen = nullptr;         /* the full-plane path (FP64, or FFT box != 128); null: LDS kernels */
    double *part_lc = nullptr;           /* lumen/ECM loop partial deviations [2][2][ORG_LUM_SPAN][ORG_LUM_MAX] */
    float *xch = nullptr;                /* tagged words exchanged by k_step's LDS loops (ORG_XCH_WORDS) */
    unsigned *lum_bar = nullptr;         /* lumen/ECM workgroups' barrier count; tag base of the LDS loops' exchanges */
    int *lum_abort = nullptr;
    long long wait_limit = 0;            /* k_step's wait timeout, wall-clock ticks */
    float *phi_win = nullptr;
    int N = 0, nc = 0, zs = 0, W = 0, M = 0;
    size_t NN = 0, P = 0;
    int grid_g = 0, ntiles = 0;
    /* full planes */
    float *phi = nullptr, *rho = nullptr, *hfunc = nullptr;
    float *lum = nullptr, *ecm = nullptr, *psi = nullptr, *hlum = nullptr, *hecm = nullptr;
    double *lbrk = nullptr;                /* lumen/ECM part of the cell update's bracket (batched step) */
    int *ringh = nullptr;                  /* ring flag of the h parts of k_mid */
    float *bufL0 = nullptr, *bufL1 = nullptr, *bufE0 = nullptr, *bufE1 = nullptr;
    float *jlx = nullptr, *jly = nullptr, *jex = nullptr, *jey = nullptr;
    /* window-local */
    float *phin_win = nullptr, *rho_box = nullptr, *vxbox = nullptr, *vybox = nullptr, *jwin = nullptr;
    float2 *spec = nullptr;
    /* per cell / scalars */
    int *xc = nullptr, *yc = nullptr, *box = nullptr, *counts = nullptr, *err = nullptr;
    int *herr = nullptr;      /* pinned host word the error bitmask is read into (a pageable one costs ~25 us) */
    int *uni = nullptr, *jbox = nullptr;   /* union of windows; per-workgroup boxes of nonzero normals */
    int *xn = nullptr, *yn = nullptr, *bn = nullptr;  /* next batched step's centres and support boxes */
    int *xp = nullptr, *yp = nullptr;      /* copy of a step's centres for k_mid */
    bool next_ready = false;               /* xn, yn, bn hold the centres of the current phi */
    int npsi = 0;                          /* psi workgroups of k_pre and k_mid */
    int nh = 0;                            /* h(lum), h(ECM) workgroups of k_pre */
    int nlp = 0;                           /* lumen/ECM copy-back workgroups of k_mid and k_post */
    int nlp_cap = 0;                       /* partial boxes allocated per field (nlp never exceeds it) */
    int nlum = ORG_LUM_BLOCKS;             /* lumen/ECM workgroups of k_step (tuning, or the batch's share) */
    int graph_steps = ORG_GRAPH_STEPS;     /* steps per replayed graph (tuning) */
    int shared_brk = 1;                    /* the bracket's shared plane (tuning) */
    int ncu = 256;                         /* compute units of the device */
    bool in_batch = false;                 /* an org_batch holds this context */
    int nparts_h = 0;                      /* parts of the sum of h(ECM): 64 x 64 tiles + the ring */
    int *live = nullptr, *live_part = nullptr;  /* live boxes of lum and ECM, k_post's partial boxes */
    bool ring_pending = true;              /* rho ghost ring not yet written since org_prepare */
    int lum_grid = ORG_LUM_BLOCKS;         /* workgroups per field of a reshaping sweep */
    double *s = nullptr, *s0 = nullptr, *sum_ecm = nullptr, *part_h = nullptr, *part_l = nullptr;
    float *f0 = nullptr;
    unsigned *counters = nullptr;
    LumState *lst = nullptr;
    bool prepared = false;
    int write_v = 1;
    int judge = ORG_JUDGE_SPECTRAL;        /* the surrogate's judge (ORG_OPT_SURROGATE_JUDGE) */
    int64_t relax_failures = 0;            /* calls in which a reshaping loop passed rimax */
    /* velocity surrogate: network input and output per cell, per-cell switch,
     * per-step judge residuals [cap][ncell] and exact-cell counts [cap] */
    float *sur_x = nullptr, *sur_v = nullptr, *sur_res = nullptr;
    int *sur_use = nullptr, *sur_nexact = nullptr;
    int sur_cap = 0;
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   /* graph_steps steps; the remainder */
    hipGraph_t graph[2] = {nullptr, nullptr};
    int graph_len[2] = {0, 0};                      /* steps captured in each */
    /* profiling */
    bool prof = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev[ORG_K_COUNT];
    size_t ev_used[ORG_K_COUNT] = {};
    std::vector<void *> allocs;
};

struct org_batch {
    std::vector<org_ctx *> ctx;
    hipStream_t stream = nullptr;
    BatchEnt *ent = nullptr;               /* one entry per organoid (device) */
    int *bmap[4] = {nullptr, nullptr, nullptr, nullptr};   /* block maps of k_pre, k_step, k_mid, k_post */
    int grid[4] = {0, 0, 0, 0};
    hipGraphExec_t gexec[2] = {nullptr, nullptr};   /* ORG_GRAPH_STEPS steps; the remainder */
    hipGraph_t graph[2] = {nullptr, nullptr};
    int graph_len[2] = {0, 0};                      /* steps captured in each */
};

namespace {

template <typename T>
int dalloc(org_ctx *c, T **ptr, size_t count) {
    void *q = nullptr;
    HIPCHK(hipMalloc(&q, (count ? count : 1) * sizeof(T)));
    c->allocs.push_back(q);
    *ptr = (T *)q;
    return ORG_OK;
}

/* Event bracket around one launch when profiling is on. */
struct Prof {
    org_ctx *c;
    int kid;
    hipStream_t st;
    hipEvent_t stop = nullptr;
    Prof(org_ctx *c_, int kid_, hipStream_t st_ = nullptr) : c(c_), kid(kid_), st(st_ ? st_ : c_->stream) {
        if (!c->prof) return;
        auto &v = c->ev[kid];
        if (c->ev_used[kid] == v.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) { c->prof = false; return; }
            v.push_back({a, b});
        }
        auto &pr = v[c->ev_used[kid]++];
        (void)hipEventRecord(pr.first, st);
        stop = pr.second;
    }
    ~Prof() {
        if (stop) (void)hipEventRecord(stop, st);
    }
};

int launch_check() {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        fprintf(stderr, "organoid: kernel launch failed: %s\n", hipGetErrorString(e));
        return note_error(ORG_EHIP);
    }
    return ORG_OK;
}

DevP make_devp(const org_params &p) {
    DevP d{};
    d.N = p.N; d.nc = p.ncell; d.zs = p.zone_size; d.W = 2 * p.zone_size; d.M = 2 * p.zone_size - 2;
    d.rimax = p.rimax;
    d.dt = (float)p.dt; d.dx = (float)p.dx; d.sigma = (float)p.sigma; d.cpl = (float)p.coupling_PF_AG;
    d.tau = (float)p.tau; d.alpha = (float)p.alpha; d.beta = (float)p.beta; d.eta = (float)p.eta;
    d.tau_lum = (float)p.tau_lum; d.beta_lc = (float)p.beta_lum_cells; d.eta_lc = (float)p.eta_lum_cells;
    d.xi_lum = (float)p.xi_lum; d.alpha_E = (float)p.alpha_ECM; d.tau_E = (float)p.tau_ECM;
    d.beta_ec = (float)p.beta_ECM_cells; d.eta_ec = (float)p.eta_ECM_cells; d.xi_E = (float)p.xi_ECM;
    d.beta_le = (float)p.beta_lum_ECM;
    d.wi = (float)p.wi; d.rdt = (float)p.rdt; d.friction = (float)p.friction; d.ncur = (float)p.n_current;
    /* label order 0 apical, 1 basal, 2/3 lateral */
    const double visc[4] = {p.visc_ap, p.visc_bas, p.visc_lat, p.visc_lat};
    const double aort[4] = {p.a_ort_ap, p.a_ort_bas, p.a_ort_lat, p.a_ort_lat};
    const double atn[4] = {p.a_tan_ap, p.a_tan_bas, p.a_tan_lat, p.a_tan_lat};
    const double bb[4] = {p.b_ap, p.b_bas, p.b_lat, p.b_lat};
    const double gen[4] = {p.gen_ap, p.gen_bas, p.gen_lat, p.gen_lat};
    const double deg[4] = {p.deg_ap, p.deg_bas, p.deg_lat, p.deg_lat};
    for (int k = 0; k < 4; k++) {
        d.visc[k] = (float)visc[k]; d.aort[k] = (float)aort[k]; d.atan_[k] = (float)atn[k];
        d.bb[k] = (float)bb[k]; d.gen[k] = (float)gen[k]; d.deg[k] = (float)deg[k];
    }
    /* stop tests as the reference evaluates them with float globals:
     * rtol*2.0*M_PI*R*wi*rdt in double (cell_phase_functions.cpp:473-475),
     * rdev_threshold*rtol*rdt in float (:720, :774) */
    const float rtol = (float)p.rtol, R = (float)p.R, wi = (float)p.wi, rdt = (float)p.rdt;
    d.thr_cell = (double)rtol * 2.0 * M_PI * (double)R * (double)wi * (double)rdt;
    d.rdev0_cell = 2.0 * (double)rtol * 2.0 * M_PI * (double)R * (double)wi * (double)rdt;
    d.thr_lum = (double)((float)p.rdev_threshold * rtol * rdt);
    d.rdev0_lum = (double)((float)p.rdev_ini * rtol * rdt);
    return d;
}

int check_ctx(org_ctx *c) {
    if (!c) return note_error(ORG_EINVAL);
    return ORG_OK;
}

int need_prepared(org_ctx *c) {
    if (!c) return note_error(ORG_EINVAL);
    c->next_ready = false;
    if (!c->prepared) {
        fprintf(stderr, "organoid: org_prepare must run after the fields are uploaded\n");
        return note_error(ORG_ESTATE);
    }
    return ORG_OK;
}

/* A reshaping loop that passed rimax is a warning, as in the reference: it
 * prints "Relaxation failure!", leaves the loop and keeps simulating
 * (cell_phase_functions.cpp:565-568, 862-865).  Counted per context. */
void note_relax_failure(org_ctx *c) {
    c->relax_failures++;
    fprintf(stderr, "organoid: Relaxation failure! (a reshaping loop passed rimax = %d; %lld so far)\n",
            c->p.rimax, (long long)c->relax_failures);
}

/* Read the device error bitmask, clear it for the next call, and map the
 * lowest set code other than ORG_ERELAX (a warning, see above). */
int harvest_device_error(org_ctx *c) {
    HIPCHK(hipMemcpyAsync(c->herr, c->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    int bits = *c->herr;
    if (!bits) return ORG_OK;
    HIPCHK(hipMemsetAsync(c->err, 0, sizeof(int), c->stream));
    if (bits & (1 << ORG_ERELAX)) {
        note_relax_failure(c);
        bits &= ~(1 << ORG_ERELAX);
    }
    for (int code = 1; code < 32; code++)
        if (bits & (1 << code)) {
            fprintf(stderr, "organoid: device reported: %s\n", org_error_string(code));
            return note_error(code);
        }
    return ORG_OK;
}

/* ---- stage launchers ------------------------------------------------------ */
int L_centroid(org_ctx *c) {
    Prof pr(c, ORG_K_CENTRES);
    hipLaunchKernelGGL(k_centroid, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->xc, c->yc, c->box, c->xc,
                       c->yc, c->box, c->uni, c->err);
    return launch_check();
}
/* the lumen/ECM state of the batched step from the current fields (see
 * k_live_init) */
int L_live_init(org_ctx *c) {
    hipLaunchKernelGGL(k_live_reset, dim3(1), dim3(256), 0, c->stream, c->live, c->live_part, c->nlp, c->N,
                       c->ringh);
    TRY(launch_check());
    hipLaunchKernelGGL(k_live_init, dim3(c->nparts_h), dim3(1024), 0, c->stream, c->dp, c->lum, c->ecm, c->hlum,
                       c->hecm, c->part_h, c->bufL0, c->bufL1, c->bufE0, c->bufE1, c->jlx, c->jly, c->jex, c->jey,
                       c->live, c->lbrk);
    return launch_check();
}
/* centres and support boxes for the next batched step, from the current
 * phi (what k_post computes at the end of every batched step) */
int L_centroid_next(org_ctx *c) {
    Prof pr(c, ORG_K_CENTRES);
    hipLaunchKernelGGL(k_centroid, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->xc, c->yc, c->box, c->xn,
                       c->yn, c->bn, c->uni, c->err);
    return launch_check();
}
int L_prep(org_ctx *c, int mode) {
    Prof pr(c, ORG_K_PREP);
    hipLaunchKernelGGL(k_prep, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->hfunc, c->xc, c->yc, c->box,
                       c->s0, c->s, c->f0, c->uni, mode);
    return launch_check();
}
int L_hlum(org_ctx *c) {
    Prof pr(c, ORG_K_LUMECM);
    hipLaunchKernelGGL(k_hlum, dim3(c->nparts_h), dim3(1024), 0, c->stream, c->dp, c->lum, c->ecm, c->hlum, c->hecm,
                       c->part_h, c->counters + 2, c->sum_ecm);
    return launch_check();
}
int L_psi(org_ctx *c) {
    Prof pr(c, ORG_K_PSI);
    hipLaunchKernelGGL(k_psi, dim3(c->ntiles), dim3(256), 0, c->stream, c->dp, c->hfunc, c->box, c->uni, c->psi);
    return launch_check();
}
int L_lumecm_update(org_ctx *c) {
    Prof pr(c, ORG_K_LUMECM);
    hipLaunchKernelGGL(k_lumecm_update, dim3(c->grid_g), dim3(1024), 0, c->stream, c->dp, c->lum, c->ecm, c->psi,
                       c->hlum, c->hecm, c->sum_ecm, c->bufL0, c->bufL1, c->bufE0, c->bufE1, c->lst, c->lum_abort,
                       c->lum_bar);
    return launch_check();
}
int L_lum_normals(org_ctx *c, int fbase, int nf) {
    Prof pr(c, ORG_K_LUMRESHAPE);
    hipLaunchKernelGGL(k_lum_normals, dim3(c->grid_g, nf), dim3(1024), 0, c->stream, c->dp, c->lum, c->ecm, c->bufL0,
                       c->bufE0, c->jlx, c->jly, c->jex, c->jey, c->jbox, fbase);
    return launch_check();
}
int L_lum_iter(org_ctx *c, int fbase, int nf) {
    Prof pr(c, ORG_K_LUMRESHAPE);
    hipLaunchKernelGGL(k_lum_iter, dim3(c->lum_grid, nf), dim3(1024), 0, c->stream, c->dp, c->bufL0, c->bufL1, c->bufE0,
                       c->bufE1, c->jlx, c->jly, c->jex, c->jey, c->jbox, c->grid_g, c->lst, c->part_l, c->counters,
                       c->err, fbase);
    return launch_check();
}
int L_commit_lumecm(org_ctx *c) {
    Prof pr(c, ORG_K_LUMRESHAPE);
    hipLaunchKernelGGL(k_commit_lumecm, dim3(c->grid_g), dim3(1024), 0, c->stream, c->dp, c->lum, c->ecm, c->bufL0,
                       c->bufL1, c->bufE0, c->bufE1, c->lst, c->err, c->uni);
    return launch_check();
}
int L_spectral(org_ctx *c) {
    Prof pr(c, ORG_K_SPECTRAL);
    hipLaunchKernelGGL(k_spectral, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->rho, c->lum, c->ecm,
                       c->xc, c->yc, c->spec, c->rho_box, c->vxbox, c->vybox, c->write_v);
    return launch_check();
}
int L_cells(org_ctx *c, int mode) {
    Prof pr(c, ORG_K_CELLS);
    hipLaunchKernelGGL(k_cells, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->rho, c->hfunc, c->psi,
                       c->hlum, c->hecm, c->f0, c->xc, c->yc, c->phin_win, c->jwin, c->counts, c->err, mode);
    return launch_check();
}
int L_commit_cells(org_ctx *c) {
    Prof pr(c, ORG_K_COMMIT);
    hipLaunchKernelGGL(k_commit_cells, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->rho, c->rho_box, c->phi,
                       (const float *)nullptr, c->xc, c->yc);
    return launch_check();
}
StepArgs step_args(org_ctx *c) {
    StepArgs a{};
    a.phi = c->phi; a.rho = c->rho; a.hfunc = c->hfunc;
    a.xc = c->xc; a.yc = c->yc; a.box = c->box; a.uni = c->uni; a.counts = c->counts;
    a.xn = c->xn; a.yn = c->yn; a.bn = c->bn; a.npsi = c->npsi;
    a.xp = c->xp; a.yp = c->yp;
    a.s0 = c->s0; a.s = c->s; a.f0 = c->f0;
    a.lum = c->lum; a.ecm = c->ecm; a.psi = c->psi; a.hlum = c->hlum; a.hecm = c->hecm;
    a.lbrk = c->lbrk; a.ringh = c->ringh; a.shared_brk = c->shared_brk;
    a.part_h = c->part_h; a.sum_ecm = c->sum_ecm; a.nh = c->nh;
    a.live = c->live; a.live_part = c->live_part; a.nlp = c->nlp; a.nlum = c->nlum;
    a.L0 = c->bufL0; a.L1 = c->bufL1; a.E0 = c->bufE0; a.E1 = c->bufE1;
    a.jlx = c->jlx; a.jly = c->jly; a.jex = c->jex; a.jey = c->jey;
    a.jbox_part = c->jbox; a.lst = c->lst; a.part_lc = c->part_lc; a.xch = c->xch;
    a.spec = c->spec; a.rho_box = c->rho_box; a.vxbox = c->vxbox; a.vybox = c->vybox; a.write_v = c->write_v;
    a.phin_win = c->phin_win; a.phi_win = c->phi_win; a.jwin = c->jwin;
    a.bar = c->lum_bar; a.abort_flag = c->lum_abort; a.err = c->err; a.limit = c->wait_limit;
    return a;
}
/* centroids, hfunc, areas and h(lum), h(ECM) of the step (see k_pre) */
int L_pre(org_ctx *c) {
    Prof pr(c, ORG_K_CENTRES);
    hipLaunchKernelGGL(k_pre, dim3(c->nc + c->npsi + c->nh), dim3(1024), 0, c->stream, c->dp, step_args(c));
    return launch_check();
}
/* the rest of the step up to the copy-back, one launch (see k_step) */
int L_step(org_ctx *c) {
    Prof pr(c, ORG_K_STEP);
    hipLaunchKernelGGL(k_step, dim3(2 * c->nc + c->nlum), dim3(1024), 0, c->stream, c->dp, step_args(c));
    return launch_check();
}
/* copy-back of one step and the first part of the next (see k_mid) */
/* k_step with the surrogate velocities of the cells whose switch is set */
int L_step_v(org_ctx *c, const float *vg, const int *use) {
    Prof pr(c, ORG_K_STEP);
    StepArgs a = step_args(c);
    a.vgiven = vg;
    a.vuse = use;
    hipLaunchKernelGGL(k_step, dim3(2 * c->nc + c->nlum), dim3(1024), 0, c->stream, c->dp, a);
    return launch_check();
}
int L_mid(org_ctx *c) {
    Prof pr(c, ORG_K_MID);
    hipLaunchKernelGGL(k_mid, dim3(c->nc + c->npsi + c->nlp), dim3(1024), 0, c->stream, c->dp, step_args(c));
    return launch_check();
}
int L_post(org_ctx *c) {
    Prof pr(c, ORG_K_COMMIT);
    hipLaunchKernelGGL(k_post, dim3(c->nc + c->nlp), dim3(1024), 0, c->stream, c->dp, step_args(c));
    return launch_check();
}
/* the one-time rho ghost ring (see k_ring_rho) */
int L_ring_if_pending(org_ctx *c) {
    if (!c->ring_pending) return ORG_OK;
    hipLaunchKernelGGL(k_ring_rho, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->rho);
    TRY(launch_check());
    c->ring_pending = false;
    return ORG_OK;
}
/* Host-driven lumen/ECM loop: run sweeps until both stop tests passed,
 * checking after every `chunk` sweeps (the reference reads rdev back every
 * sweep, cell_phase_functions.cpp:848-869). */
int run_lum_loop(org_ctx *c, int fbase, int nf, int chunk) {
    LumState st[2];
    for (long long it = 0;; it += chunk) {
        for (int k = 0; k < chunk; k++) TRY(L_lum_iter(c, fbase, nf));
        HIPCHK(hipMemcpyAsync(st, c->lst, sizeof(st), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        bool done = true;
        for (int f = fbase; f < fbase + nf; f++) done = done && st[f].done;
        if (done) return ORG_OK;
        if (it > (long long)c->p.rimax + 4) { note_relax_failure(c); return ORG_OK; }
    }
}

/* nsteps timesteps, run_with_active_gels.cpp:118-184: hfunc, areas, psi
 * and h(lum), h(ECM) of the first step (:123-142), then per step the rest up
 * to the copy-back (:130-172) followed by the copy-back merged with the next
 * step's first part, or by the plain copy-back (:172-182) after the last.
 * No host round trip: the sequence can be captured in one hipGraph. */
int enqueue_steps(org_ctx *c, int nsteps) {
    if (nsteps <= 0) return ORG_OK;
    TRY(L_pre(c));
    for (int s = 0; s < nsteps; s++) {
        TRY(L_step(c));
        TRY(s + 1 < nsteps ? L_mid(c) : L_post(c));
    }
    return ORG_OK;
}

/* ---- batches ------------------------------------------------------------ */
#define ORG_BATCH_MAX 255
enum { BK_PRE = 0, BK_STEP = 1, BK_MID = 2, BK_POST = 3 };

int LB(org_batch *b, int k) {
    const ORG_CONST BatchEnt *ent = (const ORG_CONST BatchEnt *)b->ent;
    const ORG_CONST int *map = (const ORG_CONST int *)b->bmap[k];
    switch (k) {
    case BK_PRE: hipLaunchKernelGGL(k_pre_batch, dim3(b->grid[k]), dim3(1024), 0, b->stream, ent, map); break;
    case BK_STEP: hipLaunchKernelGGL(k_step_batch, dim3(b->grid[k]), dim3(1024), 0, b->stream, ent, map); break;
    case BK_MID: hipLaunchKernelGGL(k_mid_batch, dim3(b->grid[k]), dim3(1024), 0, b->stream, ent, map); break;
    default: hipLaunchKernelGGL(k_post_batch, dim3(b->grid[k]), dim3(1024), 0, b->stream, ent, map); break;
    }
    return launch_check();
}

/* the stage sequence of enqueue_steps over every organoid of the batch */
int enqueue_batch_steps(org_batch *b, int nsteps) {
    if (nsteps <= 0) return ORG_OK;
    TRY(LB(b, BK_PRE));
    for (int s = 0; s < nsteps; s++) {
        TRY(LB(b, BK_STEP));
        TRY(LB(b, s + 1 < nsteps ? BK_MID : BK_POST));
    }
    return ORG_OK;
}

int batch_upload_entries(org_batch *b) {
    std::vector<BatchEnt> h(b->ctx.size());
    for (size_t o = 0; o < b->ctx.size(); o++) { h[o].P = b->ctx[o]->dp; h[o].a = step_args(b->ctx[o]); }
    HIPCHK(hipMemcpyAsync(b->ent, h.data(), sizeof(BatchEnt) * h.size(), hipMemcpyHostToDevice, b->stream));
    HIPCHK(hipStreamSynchronize(b->stream));
    return ORG_OK;
}

void batch_drop_graphs(org_batch *b) {
    for (int k = 0; k < 2; k++) {
        if (b->gexec[k]) { (void)hipGraphExecDestroy(b->gexec[k]); b->gexec[k] = nullptr; }
        if (b->graph[k]) { (void)hipGraphDestroy(b->graph[k]); b->graph[k] = nullptr; }
    }
}

/* The replayable graph of `len` steps in slot (g, ge, have): captured from
 * `enqueue(len)` on stream st, or reused when the slot already holds that
 * length.  A replay of n steps is (n / L) graphs of L steps and one graph of
 * the n % L remaining steps, so every replay runs k_step + k_mid between
 * steps and k_pre / k_post only at its two ends. */
template <typename F>
int ensure_graph(hipStream_t st, hipGraph_t &g, hipGraphExec_t &ge, int &have, int len, F enqueue) {
    if (ge && have == len) return ORG_OK;
    if (ge) { (void)hipGraphExecDestroy(ge); ge = nullptr; }
    if (g) { (void)hipGraphDestroy(g); g = nullptr; }
    have = 0;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(len);
    hipGraph_t cap = nullptr;
    hipError_t e = hipStreamEndCapture(st, &cap);
    if (rc || e != hipSuccess) {
        if (cap) (void)hipGraphDestroy(cap);
        return rc ? rc : note_error(ORG_EHIP);
    }
    g = cap;
    HIPCHK(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    have = len;
    return ORG_OK;
}

int batch_harvest(org_batch *b) {
    HIPCHK(hipStreamSynchronize(b->stream));
    int rc = ORG_OK;
    for (org_ctx *c : b->ctx) {
        const int r = harvest_device_error(c);
        if (r && !rc) rc = r;
    }
    return rc;
}

}  // namespace

/* ======================================================================== */
extern "C" {

int org_last_error(void) { return g_last_error.load(); }

int org_get_relax_failures(org_ctx *c, int64_t *count) {
    TRY(check_ctx(c));
    if (!count) return note_error(ORG_EINVAL);
    *count = c->relax_failures;
    return ORG_OK;
}

const char *org_error_string(int code) {
    switch (code) {
    case ORG_OK: return "ok";
    case ORG_EINVAL: return "invalid argument";
    case ORG_EHIP: return "HIP runtime error";
    case ORG_EEMPTY_CELL: return "a cell has no phi>0.5 pixel (centroid undefined)";
    case ORG_EWINDOW: return "a cell window (+1 halo) leaves the grid interior";
    case ORG_ERELAX: return "reshaping did not converge (rimax)";
    case ORG_ENOTBUILT: return "kernels not built for this device";
    case ORG_ESTATE: return "call order / state error";
    case ORG_ETIMEOUT: return "a wait between workgroups of one launch timed out";
    default: return "unknown error";
    }
}

const char *org_build_info(void) { return "organoid_hip gfx950"; }

/* A context of the full-plane path: FP64 (f64 = 1), or FP32 with an FFT box
 * other than the LDS kernels' 128. */
static int create_generic_ctx(const org_params *p, const double *s0_cell, int f64, org_ctx **out) {
    if (p->ncell < 2 || p->ncell > 1024 || p->zone_size < 3 || p->zone_size > 513 || p->N < 2 * p->zone_size + 6 ||
        p->n_current <= 0) {
        fprintf(stderr, "organoid: need 2 <= ncell <= 1024, 3 <= zone_size <= 513, N >= 2*zone_size+6, "
                        "n_current > 0\n");
        return note_error(ORG_EINVAL);
    }
    org_ctx *c = new org_ctx();
    c->p = *p;
    c->N = p->N; c->nc = p->ncell; c->zs = p->zone_size; c->W = 2 * c->zs; c->M = 2 * c->zs - 2;
    c->NN = (size_t)c->N * c->N; c->P = c->NN * c->nc;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        org_destroy(c);
        return note_error(ORG_EHIP);
    }
    const int rc = gen::create_generic(p, s0_cell, f64, c->stream, &c->gen);
    if (rc) {
        org_destroy(c);
        return note_error(rc);
    }
    *out = c;
    return ORG_OK;
}

int org_create_f64(const org_params *p, const double *s0_cell, org_ctx **out) {
    if (!p || !out) return note_error(ORG_EINVAL);
    *out = nullptr;
    return create_generic_ctx(p, s0_cell, 1, out);
}

int org_get_precision(org_ctx *c, int32_t *real_bytes) {
    TRY(check_ctx(c));
    if (!real_bytes) return note_error(ORG_EINVAL);
    *real_bytes = c->gen ? c->gen->real_bytes() : (int32_t)sizeof(float);
    return ORG_OK;
}

int org_create(const org_params *p, const double *s0_cell, org_ctx **out) {
    if (!p || !out) return note_error(ORG_EINVAL);
    *out = nullptr;
    if (p->zone_size != ORG_ZS_MAX) return create_generic_ctx(p, s0_cell, 0, out);
    if (p->ncell < 2 || p->ncell > 1024 || p->N < 2 * p->zone_size + 6 || p->n_current <= 0) {
        fprintf(stderr, "organoid: need 2 <= ncell <= 1024, N >= 2*zone_size+6, n_current > 0\n");
        return note_error(ORG_EINVAL);
    }
    org_ctx *c = new org_ctx();
    c->p = *p;
    c->dp = make_devp(*p);
    c->N = p->N; c->nc = p->ncell; c->zs = p->zone_size; c->W = 2 * c->zs; c->M = 2 * c->zs - 2;
    c->NN = (size_t)c->N * c->N; c->P = c->NN * c->nc;
    c->grid_g = std::min(c->N, 128);
    {
        const int t = (c->N + 63) / 64;
        c->nparts_h = t * t + 1;
        /* k_pre and k_mid: the cells' workgroups plus two equal roles that
         * together fill the remaining CUs (one 1024-thread workgroup per CU
         * starts at once; more wait for a free CU) */
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
            ncu = 256;
        const int role = std::max(32, std::max(64, ncu - c->nc) / 2);
        c->ncu = ncu;
        c->npsi = role;
        c->nh = std::min(c->nparts_h, role);
        c->nlp = role;
        c->nlp_cap = std::max(role, ORG_ROLE_MAX);
    }
    int nt = (c->N + 15) / 16;
    c->ntiles = nt * nt;
    int rc = ORG_OK;
    auto A = [&](auto **ptr, size_t n) { if (!rc) rc = dalloc(c, ptr, n); };
    A(&c->phi, c->P); A(&c->rho, c->P); A(&c->hfunc, c->P);
    A(&c->lum, c->NN); A(&c->ecm, c->NN); A(&c->psi, c->NN); A(&c->hlum, c->NN); A(&c->hecm, c->NN);
    A(&c->lbrk, c->NN); A(&c->ringh, 1);
    A(&c->bufL0, c->NN); A(&c->bufL1, c->NN); A(&c->bufE0, c->NN); A(&c->bufE1, c->NN);
    A(&c->jlx, c->NN); A(&c->jly, c->NN); A(&c->jex, c->NN); A(&c->jey, c->NN);
    A(&c->phin_win, (size_t)c->nc * c->W * c->W);
    A(&c->rho_box, (size_t)c->nc * c->M * c->M);
    A(&c->vxbox, (size_t)c->nc * c->M * c->M);
    A(&c->vybox, (size_t)c->nc * c->M * c->M);
    A(&c->jwin, (size_t)c->nc * 2 * (c->W + 2) * (c->W + 2));
    A(&c->spec, (size_t)c->nc * 2 * c->M * c->M);
    A(&c->uni, 4); A(&c->jbox, 2 * 4 * 1024); A(&c->part_lc, 4 * ORG_LUM_SPAN * ORG_LUM_MAX); A(&c->xch, 2 * ORG_XCH_WORDS(c->NN)); A(&c->lum_bar, 2); A(&c->lum_abort, 1);
    A(&c->phi_win, (size_t)c->nc * c->W * c->W);
    A(&c->xc, c->nc); A(&c->yc, c->nc); A(&c->box, 4 * (size_t)c->nc);
    A(&c->xn, c->nc); A(&c->yn, c->nc); A(&c->bn, 4 * (size_t)c->nc);
    A(&c->xp, c->nc); A(&c->yp, c->nc);
    A(&c->live, 8); A(&c->live_part, 8 * (size_t)c->nlp_cap); A(&c->counts, c->nc); A(&c->err, 1);
    A(&c->s, c->nc); A(&c->s0, c->nc); A(&c->sum_ecm, 1); A(&c->part_h, (size_t)c->nparts_h); A(&c->part_l, 2 * 2048);
    A(&c->f0, c->nc); A(&c->counters, 4); A(&c->lst, 2);
    if (!rc && hipHostMalloc((void **)&c->herr, sizeof(int)) != hipSuccess) rc = note_error(ORG_EHIP);
    if (!rc && hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = note_error(ORG_EHIP);
    if (!rc) {
        /* timeout of the waits inside k_step: 2 s of the device's wall clock */
        int dev = 0, khz = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0)
            rc = note_error(ORG_EHIP);
        else
            c->wait_limit = 2000LL * khz;
    }
    if (rc) { org_destroy(c); return rc; }
    std::vector<double> s0v(c->nc);
    for (int n = 0; n < c->nc; n++) s0v[n] = (p->use_s0_cell && s0_cell) ? s0_cell[n] : p->s0;
    std::vector<int> cen(c->nc, c->N / 2);
    if (hipMemcpy(c->s0, s0v.data(), sizeof(double) * c->nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->xc, cen.data(), sizeof(int) * c->nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->yc, cen.data(), sizeof(int) * c->nc, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->err, 0, sizeof(int)) != hipSuccess || hipMemset(c->counters, 0, 4 * sizeof(unsigned)) != hipSuccess ||
        hipMemset(c->counts, 0, sizeof(int) * c->nc) != hipSuccess || hipMemset(c->lst, 0, 2 * sizeof(LumState)) != hipSuccess ||
        hipMemset(c->hfunc, 0, sizeof(float) * c->P) != hipSuccess ||
        hipMemset(c->hlum, 0, sizeof(float) * c->NN) != hipSuccess ||
        hipMemset(c->hecm, 0, sizeof(float) * c->NN) != hipSuccess ||
        hipMemset(c->sum_ecm, 0, sizeof(double)) != hipSuccess ||
        hipMemset(c->lum_bar, 0, 2 * sizeof(unsigned)) != hipSuccess || hipMemset(c->lum_abort, 0, sizeof(int)) != hipSuccess ||
        hipMemset(c->xch, 0, sizeof(unsigned long long) * ORG_XCH_WORDS(c->NN)) != hipSuccess ||
        hipMemset(c->ringh, 0, sizeof(int)) != hipSuccess ||
        /* the loops read normals one pixel outside the live box, where none
         * is ever computed in box mode: they must be zero */
        hipMemset(c->jlx, 0, sizeof(float) * c->NN) != hipSuccess || hipMemset(c->jly, 0, sizeof(float) * c->NN) != hipSuccess ||
        hipMemset(c->jex, 0, sizeof(float) * c->NN) != hipSuccess || hipMemset(c->jey, 0, sizeof(float) * c->NN) != hipSuccess) {
        org_destroy(c);
        return note_error(ORG_EHIP);
    }
    *out = c;
    return ORG_OK;
}

int org_destroy(org_ctx *c) {
    if (!c) return ORG_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->gen;
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) (void)hipGraphExecDestroy(c->gexec[k]);
        if (c->graph[k]) (void)hipGraphDestroy(c->graph[k]);
    }
    for (int k = 0; k < ORG_K_COUNT; k++)
        for (auto &pr : c->ev[k]) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
    for (void *q : c->allocs) (void)hipFree(q);
    if (c->herr) (void)hipHostFree(c->herr);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ORG_OK;
}

static float *field_ptr(org_ctx *c, int f, size_t *n) {
    switch (f) {
    case ORG_F_PHI: *n = c->P; return c->phi;
    case ORG_F_RHO: *n = c->P; return c->rho;
    case ORG_F_HFUNC: *n = c->P; return c->hfunc;
    case ORG_F_LUM: *n = c->NN; return c->lum;
    case ORG_F_ECM: *n = c->NN; return c->ecm;
    case ORG_F_PSI: *n = c->NN; return c->psi;
    case ORG_F_HLUM: *n = c->NN; return c->hlum;
    case ORG_F_HECM: *n = c->NN; return c->hecm;
    case ORG_F_LUMN: *n = c->NN; return c->bufL0;
    case ORG_F_ECMN: *n = c->NN; return c->bufE0;
    default: *n = 0; return nullptr;
    }
}

static int gen_rc(int rc) { return rc ? note_error(rc) : ORG_OK; }

static void unprepare_on(org_ctx *c, int field) {
    if (field == ORG_F_PHI || field == ORG_F_HFUNC || field == ORG_F_RHO) c->prepared = false;
    c->next_ready = false;
}

int org_set_field(org_ctx *c, int field, const float *host) {
    TRY(check_ctx(c));
    if (c->gen) {
        TRY(gen_rc(c->gen->set_field(field, host, 0)));
        unprepare_on(c, field);
        return ORG_OK;
    }
    size_t n;
    float *d = field_ptr(c, field, &n);
    if (!d || !host) return note_error(ORG_EINVAL);
    HIPCHK(hipMemcpyAsync(d, host, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (field == ORG_F_PHI || field == ORG_F_HFUNC || field == ORG_F_RHO) c->prepared = false;
    c->next_ready = false;
    return ORG_OK;
}

int org_get_field(org_ctx *c, int field, float *host) {
    TRY(check_ctx(c));
    if (c->gen) return gen_rc(c->gen->get_field(field, host, 0));
    size_t n;
    float *d = field_ptr(c, field, &n);
    if (!d || !host) return note_error(ORG_EINVAL);
    if (field == ORG_F_LUMN || field == ORG_F_ECMN) {
        /* the last iterate of the reshaping loop */
        LumState st[2];
        HIPCHK(hipMemcpyAsync(st, c->lst, sizeof(st), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        if (field == ORG_F_LUMN && st[0].cur) d = c->bufL1;
        if (field == ORG_F_ECMN && st[1].cur) d = c->bufE1;
    }
    HIPCHK(hipMemcpyAsync(host, d, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ORG_OK;
}

int org_get_window(org_ctx *c, int which, float *host) {
    TRY(check_ctx(c));
    if (!host) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_window(which, host, 0));
    const float *d = nullptr;
    size_t n = 0;
    switch (which) {
    case ORG_W_PHIN: d = c->phin_win; n = (size_t)c->nc * c->W * c->W; break;
    case ORG_W_RHON: d = c->rho_box; n = (size_t)c->nc * c->M * c->M; break;
    case ORG_W_VX: d = c->vxbox; n = (size_t)c->nc * c->M * c->M; break;
    case ORG_W_VY: d = c->vybox; n = (size_t)c->nc * c->M * c->M; break;
    default: return note_error(ORG_EINVAL);
    }
    HIPCHK(hipMemcpyAsync(host, d, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ORG_OK;
}

int org_get_centres(org_ctx *c, int32_t *xc, int32_t *yc) {
    TRY(check_ctx(c));
    if (!xc || !yc) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_centres(xc, yc));
    HIPCHK(hipMemcpyAsync(xc, c->xc, sizeof(int) * c->nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(yc, c->yc, sizeof(int) * c->nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ORG_OK;
}

int org_get_areas(org_ctx *c, double *s, double *f0) {
    TRY(check_ctx(c));
    if (!s || !f0) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_areas(s, f0));
    std::vector<float> f(c->nc);
    HIPCHK(hipMemcpyAsync(s, c->s, sizeof(double) * c->nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(f.data(), c->f0, sizeof(float) * c->nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int n = 0; n < c->nc; n++) f0[n] = f[n];
    return ORG_OK;
}

int org_get_iterations(org_ctx *c, int32_t *cells, int32_t *lum, int32_t *ecm) {
    TRY(check_ctx(c));
    if (!cells || !lum || !ecm) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_iterations(cells, lum, ecm));
    LumState st[2];
    HIPCHK(hipMemcpyAsync(cells, c->counts, sizeof(int) * c->nc, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(st, c->lst, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    *lum = st[0].ri; *ecm = st[1].ri;
    return ORG_OK;
}

int org_prepare(org_ctx *c) {
    TRY(check_ctx(c));
    if (c->gen) {
        TRY(gen_rc(c->gen->prepare()));
        c->prepared = true;
        return ORG_OK;
    }
    hipLaunchKernelGGL(k_init_box, dim3(c->nc), dim3(256), 0, c->stream, c->dp, c->phi, c->hfunc, c->box);
    TRY(launch_check());
    HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(float) * c->NN, c->stream));
    HIPCHK(hipMemsetAsync(c->lum_bar, 0, sizeof(unsigned), c->stream));
    HIPCHK(hipMemsetAsync(c->lum_abort, 0, sizeof(int), c->stream));
    /* lumn <- lum, ECMn <- ECM (run_with_active_gels.cpp:92-95): both
     * ping-pong buffers of each field start as the uploaded field, and the
     * loop state points at the first, so ORG_F_LUMN / ORG_F_ECMN read before
     * the first update return the uploaded fields */
    HIPCHK(hipMemsetAsync(c->lst, 0, 2 * sizeof(LumState), c->stream));
    HIPCHK(hipMemcpyAsync(c->bufL0, c->lum, sizeof(float) * c->NN, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bufL1, c->lum, sizeof(float) * c->NN, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bufE0, c->ecm, sizeof(float) * c->NN, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->bufE1, c->ecm, sizeof(float) * c->NN, hipMemcpyDeviceToDevice, c->stream));
    {
        const int u[4] = {c->N, 0, c->N, 0};
        HIPCHK(hipStreamSynchronize(c->stream));
        HIPCHK(hipMemcpy(c->uni, u, sizeof(u), hipMemcpyHostToDevice));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prepared = true;
    c->ring_pending = true;
    c->next_ready = false;
    return ORG_OK;
}

int org_get_tuning(org_ctx *c, org_tuning *t) {
    TRY(check_ctx(c));
    if (!t) return note_error(ORG_EINVAL);
    t->lum_blocks = c->nlum;
    t->roles = c->npsi;
    t->graph_steps = c->graph_steps;
    t->shared_bracket = c->shared_brk;
    return ORG_OK;
}

int org_set_tuning(org_ctx *c, const org_tuning *t) {
    TRY(check_ctx(c));
    if (!t || c->in_batch) return note_error(ORG_EINVAL);
    if (c->gen) return ORG_OK;               /* the full-plane path has no tunable launch shape */
    if (t->lum_blocks < 1 || t->lum_blocks > std::min(ORG_LUM_MAX, c->ncu) || t->roles < 1 ||
        t->roles > c->nlp_cap || t->graph_steps < 1 || t->graph_steps > 64 ||
        (t->shared_bracket != 0 && t->shared_bracket != 1))
        return note_error(ORG_EINVAL);
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nlum = t->lum_blocks;
    c->npsi = t->roles;
    c->nh = std::min(c->nparts_h, t->roles);
    if (c->nlp != t->roles) c->next_ready = false;   /* partial live boxes are re-initialised for the new nlp */
    c->nlp = t->roles;
    c->graph_steps = t->graph_steps;
    /* the shared plane is not kept up to date without it: k_live_init
     * recomputes it before the next batched step */
    if (t->shared_bracket && !c->shared_brk) c->next_ready = false;
    c->shared_brk = t->shared_bracket;
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) { (void)hipGraphExecDestroy(c->gexec[k]); c->gexec[k] = nullptr; }
        if (c->graph[k]) { (void)hipGraphDestroy(c->graph[k]); c->graph[k] = nullptr; }
    }
    return ORG_OK;
}

int org_set_option(org_ctx *c, int option, int value) {
    TRY(check_ctx(c));
    if (c->gen) return option == ORG_OPT_WRITE_VELOCITY || option == ORG_OPT_SURROGATE_JUDGE ? ORG_OK
                                                                                              : note_error(ORG_EINVAL);
    switch (option) {
    case ORG_OPT_WRITE_VELOCITY: c->write_v = value ? 1 : 0; break;
    case ORG_OPT_SURROGATE_JUDGE:
        if (value != ORG_JUDGE_SPECTRAL && value != ORG_JUDGE_FINITE_DIFF) return note_error(ORG_EINVAL);
        c->judge = value;
        return ORG_OK;        /* the batched step's graphs do not use it */

    default: return note_error(ORG_EINVAL);
    }
    for (int k = 0; k < 2; k++) {
        if (c->gexec[k]) { (void)hipGraphExecDestroy(c->gexec[k]); c->gexec[k] = nullptr; }
        if (c->graph[k]) { (void)hipGraphDestroy(c->graph[k]); c->graph[k] = nullptr; }
    }
    return ORG_OK;
}

/* ---- per-function entry points -----------------------------------------
 * Each may change phi or the centres, so the next batched step recomputes
 * its centres first. */
#define GEN_ROUTE(call) do { if (c->gen) return gen_rc(c->gen->call); } while (0)
int org_find_centres_cells(org_ctx *c) {
    TRY(need_prepared(c));
    GEN_ROUTE(find_centres());
    TRY(L_centroid(c));
    return harvest_device_error(c);
}
int org_compute_hfunc_cells(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(hfunc_cells()); return L_prep(c, 1); }
int org_sum_s(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(sum_s()); return L_prep(c, 2); }
int org_compute_hfunc_lum_ecm(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(hfunc_lum_ecm()); return L_hlum(c); }
int org_compute_psi(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(psi()); return L_psi(c); }
int org_update_all_cortices(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(cortices()); return L_spectral(c); }
int org_update_lumen_and_ECM(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(lumecm_update()); return L_lumecm_update(c); }
/* the full-plane path's reshaping calls, with the rimax warning */
static int gen_reshape(org_ctx *c, int what) {
    int failed = 0;
    const int rc = what == 2 ? c->gen->reshape_cells(0, &failed)
                 : what == 3 ? c->gen->reshape_cells(1, &failed)
                             : c->gen->reshape_lumecm(what, &failed);
    if (failed) note_relax_failure(c);
    return gen_rc(rc);
}
int org_reshaping_lumen_or_ECM(org_ctx *c, int which_ecm) {
    TRY(need_prepared(c));
    if (c->gen) return gen_reshape(c, which_ecm ? 1 : 0);
    int f = which_ecm ? 1 : 0;
    TRY(L_lum_normals(c, f, 1));
    TRY(run_lum_loop(c, f, 1, 1));
    TRY(harvest_device_error(c));
    /* the reference's loop copies lum <- lumn before every sweep
     * (cell_phase_functions.cpp:777-785): its first argument ends at the
     * second-to-last iterate, the other buffer of the pair */
    LumState st[2];
    HIPCHK(hipMemcpyAsync(st, c->lst, sizeof(st), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st[f].ri >= 1) {
        const float *prev = f ? (st[f].cur ? c->bufE0 : c->bufE1) : (st[f].cur ? c->bufL0 : c->bufL1);
        HIPCHK(hipMemcpyAsync(f ? c->ecm : c->lum, prev, sizeof(float) * c->NN, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    return ORG_OK;
}
int org_update_all_cells(org_ctx *c) { TRY(need_prepared(c)); GEN_ROUTE(update_cells()); return L_cells(c, 1); }
int org_reshaping_cells(org_ctx *c) {
    TRY(need_prepared(c));
    if (c->gen) return gen_reshape(c, 2);
    TRY(L_cells(c, 2));
    return harvest_device_error(c);
}
int org_reshaping_cells_from_phi(org_ctx *c) {
    TRY(need_prepared(c));
    if (c->gen) return gen_reshape(c, 3);
    TRY(L_cells(c, 2 | 4));
    return harvest_device_error(c);
}
int org_update_rho_to_shape(org_ctx *c, float rho0, float tol) {
    TRY(need_prepared(c));
    GEN_ROUTE(rho_to_shape(rho0, tol));
    hipLaunchKernelGGL(k_rho_to_shape, dim3(1024), dim3(256), 0, c->stream, c->phi, c->rho, c->P, rho0, tol);
    TRY(launch_check());
    HIPCHK(hipStreamSynchronize(c->stream));
    return ORG_OK;
}
/* The IC program's device part (prepare_initial_files.cpp:82-110): centres,
 * reshaping_cells from phin = phi, update_rho_to_shape, update_lumen_and_ECM
 * and both reshaping loops.  The program never computes psi, h(lum), h(ECM)
 * (created, not copied in, :80): they and the sum of h(ECM) are taken as
 * zero.  No copy-back follows, so lum and ECM end at their second-to-last
 * iterates and phi at its second-to-last reshaping iterate, as in the
 * program's saved frame (:108). */
int org_update_rho_to_shape_f64(org_ctx *c, double rho0, double tol) {
    TRY(need_prepared(c));
    if (c->gen) return gen_rc(c->gen->rho_to_shape(rho0, tol));
    return org_update_rho_to_shape(c, (float)rho0, (float)tol);
}
int org_relax_initial_state(org_ctx *c, float rho0, float tol) {
    TRY(need_prepared(c));
    if (c->gen) {
        TRY(gen_rc(c->gen->find_centres()));
        TRY(gen_reshape(c, 3));
        TRY(gen_rc(c->gen->rho_to_shape(rho0, tol)));
        TRY(gen_rc(c->gen->zero_gel_inputs()));
        TRY(gen_rc(c->gen->lumecm_update()));
        TRY(gen_reshape(c, 0));
        return gen_reshape(c, 1);
    }
    TRY(L_centroid(c));
    TRY(harvest_device_error(c));
    TRY(org_reshaping_cells_from_phi(c));
    TRY(org_update_rho_to_shape(c, rho0, tol));
    HIPCHK(hipMemsetAsync(c->psi, 0, sizeof(float) * c->NN, c->stream));
    HIPCHK(hipMemsetAsync(c->hlum, 0, sizeof(float) * c->NN, c->stream));
    HIPCHK(hipMemsetAsync(c->hecm, 0, sizeof(float) * c->NN, c->stream));
    HIPCHK(hipMemsetAsync(c->sum_ecm, 0, sizeof(double), c->stream));
    TRY(org_update_lumen_and_ECM(c));
    TRY(org_reshaping_lumen_or_ECM(c, 0));
    TRY(org_reshaping_lumen_or_ECM(c, 1));
    c->next_ready = false;
    return harvest_device_error(c);
}
int org_commit_step(org_ctx *c) {
    TRY(need_prepared(c));
    if (c->gen) {
        TRY(gen_rc(c->gen->commit_rho()));
        return gen_rc(c->gen->commit_lumecm());
    }
    TRY(L_commit_cells(c));
    TRY(L_ring_if_pending(c));
    TRY(L_commit_lumecm(c));
    return harvest_device_error(c);
}
int org_commit_rho(org_ctx *c) {
    TRY(need_prepared(c));
    GEN_ROUTE(commit_rho());
    TRY(L_commit_cells(c));
    return L_ring_if_pending(c);
}
int org_commit_lumecm(org_ctx *c) {
    TRY(need_prepared(c));
    GEN_ROUTE(commit_lumecm());
    TRY(L_commit_lumecm(c));
    return harvest_device_error(c);
}
#undef GEN_ROUTE

/* ---- batched form ------------------------------------------------------- */
/* One timestep of the full-plane path in the driver's order
 * (run_with_active_gels.cpp:118-184). */
static int gen_step_body(org_ctx *c) {
    gen::Generic *g = c->gen;
    TRY(gen_rc(g->find_centres()));                      /* :123 */
    TRY(gen_rc(g->hfunc_cells()));                       /* :133 */
    TRY(gen_rc(g->hfunc_lum_ecm()));                     /* :134 */
    TRY(gen_rc(g->psi()));                               /* :139 */
    TRY(gen_rc(g->sum_s()));                             /* :142 */
    TRY(gen_rc(g->cortices()));                          /* :130-154 */
    TRY(gen_rc(g->lumecm_update()));                     /* :157-158 */
    {                                                    /* :160-161 */
        int failed[2] = {0, 0};
        const int rc = g->reshape_lumecm_pair(failed);
        for (int f = 0; f < 2; f++)
            if (failed[f]) note_relax_failure(c);
        TRY(gen_rc(rc));
    }
    TRY(gen_rc(g->update_cells()));                      /* :169-170 */
    TRY(gen_reshape(c, 2));                              /* :164-166, :172 */
    g->step_scope(false);                                /* phi changes from here on */
    TRY(gen_rc(g->commit_rho()));                        /* :175-178 */
    return gen_rc(g->commit_lumecm());                   /* :181-182 */
}
static int gen_step(org_ctx *c) {
    c->gen->step_scope(true);      /* |grad phi| formed once: phi is fixed until the copy-back */
    const int rc = gen_step_body(c);
    c->gen->step_scope(false);
    return rc;
}

int org_step(org_ctx *c, int it0, int nsteps, int use_graph) {
    (void)it0;
    if (!c) return note_error(ORG_EINVAL);
    if (c->gen) {                          /* host-driven: the reshaping loops' stop tests read back */
        TRY(need_prepared(c));
        if (nsteps < 0) return note_error(ORG_EINVAL);
        int rc = ORG_OK;
        c->gen->step_chain(true);          /* centroid sums carried from step to step */
        for (int s = 0; s < nsteps && !rc; s++) rc = gen_step(c);
        c->gen->step_chain(false);
        return rc;
    }
    const bool ready = c->next_ready;
    TRY(need_prepared(c));
    if (nsteps < 0) return note_error(ORG_EINVAL);
    if (nsteps == 0) { c->next_ready = ready; return ORG_OK; }
    if (!ready) {                          /* k_post keeps both up to date after every batched step */
        TRY(L_centroid_next(c));
        TRY(L_live_init(c));
    }
    c->next_ready = true;
    if (nsteps > 0 && c->ring_pending) {   /* first step after org_prepare: eager, plus the ghost ring */
        TRY(enqueue_steps(c, 1));
        TRY(L_ring_if_pending(c));
        nsteps--;
    }
    if (!use_graph) {
        TRY(enqueue_steps(c, nsteps));
        return harvest_device_error(c);
    }
    /* graphs of graph_steps steps and of the remaining nsteps % graph_steps */
    const bool prof = c->prof;
    c->prof = false;                     /* no event records inside the graphs */
    auto enq = [c](int n) { return enqueue_steps(c, n); };
    const int full = nsteps / c->graph_steps, rem = nsteps % c->graph_steps;
    int rc = ORG_OK;
    if (full) rc = ensure_graph(c->stream, c->graph[0], c->gexec[0], c->graph_len[0], c->graph_steps, enq);
    if (!rc && rem) rc = ensure_graph(c->stream, c->graph[1], c->gexec[1], c->graph_len[1], rem, enq);
    c->prof = prof;
    TRY(rc);
    for (int s = 0; s < full; s++) HIPCHK(hipGraphLaunch(c->gexec[0], c->stream));
    if (rem) HIPCHK(hipGraphLaunch(c->gexec[1], c->stream));
    return harvest_device_error(c);
}

int org_graph_prepare(org_ctx *c, int nsteps) {
    if (!c || nsteps < 0) return note_error(ORG_EINVAL);
    if (c->gen) return ORG_OK;
    TRY(need_prepared(c));
    const int n = nsteps - (c->ring_pending && nsteps > 0 ? 1 : 0);   /* org_step runs that one eagerly */
    const int full = n / c->graph_steps, rem = n % c->graph_steps;
    const bool prof = c->prof;
    c->prof = false;
    auto enq = [c](int k) { return enqueue_steps(c, k); };
    int rc = ORG_OK;
    if (full) rc = ensure_graph(c->stream, c->graph[0], c->gexec[0], c->graph_len[0], c->graph_steps, enq);
    if (!rc && rem) rc = ensure_graph(c->stream, c->graph[1], c->gexec[1], c->graph_len[1], rem, enq);
    c->prof = prof;
    return rc;
}

/* ---- velocity surrogate ----------------------------------------------------- */
namespace {
/* buffers of the surrogate for nsteps steps (per-step arrays regrown) */
int sur_buffers(org_ctx *c, int nsteps) {
    const size_t MM = (size_t)ORG_MB * ORG_MB;
    if (!c->sur_x) {
        TRY(dalloc(c, &c->sur_x, (size_t)c->nc * 2 * MM));
        TRY(dalloc(c, &c->sur_v, (size_t)c->nc * 2 * MM));
        TRY(dalloc(c, &c->sur_use, (size_t)c->nc));
    }
    if (nsteps > c->sur_cap) {
        for (void *q : {(void *)c->sur_res, (void *)c->sur_nexact}) {
            if (!q) continue;
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(q));
            c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), q));
        }
        c->sur_res = nullptr; c->sur_nexact = nullptr;
        TRY(dalloc(c, &c->sur_res, (size_t)nsteps * c->nc));
        TRY(dalloc(c, &c->sur_nexact, (size_t)nsteps));
        c->sur_cap = nsteps;
    }
    return ORG_OK;
}
/* the batched step's set-up before its first launch (see org_step) */
int step_prologue(org_ctx *c) {
    if (!c->next_ready) {
        TRY(L_centroid_next(c));
        TRY(L_live_init(c));
    }
    c->next_ready = true;
    return ORG_OK;
}
/* after step s of a per-step loop: copy-back, the first step's rho ring,
 * and whether the next step needs k_pre (k_mid prepared it otherwise) */
int step_epilogue(org_ctx *c, bool last, bool &need_pre) {
    if (last || c->ring_pending) {
        TRY(L_post(c));
        TRY(L_ring_if_pending(c));
        need_pre = true;
    } else {
        TRY(L_mid(c));
        need_pre = false;
    }
    return ORG_OK;
}
}  // namespace

int org_surrogate_samples(org_ctx *c, int it0, int nsteps, float *x, float *v, float *res) {
    (void)it0;
    if (!c || !x || !v || nsteps < 1 || c->gen) return note_error(ORG_EINVAL);
    TRY(need_prepared(c));
    TRY(sur_buffers(c, 1));
    TRY(step_prologue(c));
    const size_t MM = (size_t)ORG_MB * ORG_MB, per = (size_t)c->nc * 2 * MM;
    const int wv = c->write_v;
    c->write_v = 1;
    int rc = ORG_OK;
    bool need_pre = true;
    for (int s = 0; s < nsteps && !rc; s++) {
        if (need_pre && (rc = L_pre(c))) break;
        hipLaunchKernelGGL(k_crops, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->rho, c->xc, c->yc,
                           x + s * per);
        if ((rc = launch_check()) || (rc = L_step(c))) break;
        hipLaunchKernelGGL(k_pack_v, dim3(c->nc), dim3(1024), 0, c->stream, c->vxbox, c->vybox, v + s * per);
        if ((rc = launch_check())) break;
        if (res) {        /* phi, rho, lum, ECM and the centres are still the step's own until the copy-back */
            hipLaunchKernelGGL(c->judge == ORG_JUDGE_SPECTRAL ? k_vresid_spec : k_vresid, dim3(c->nc), dim3(1024), 0,
                               c->stream, c->dp, c->phi, c->rho, c->lum, c->ecm, c->xc, c->yc, c->spec,
                               v + s * per, INFINITY, res + (size_t)s * c->nc, c->sur_use, c->sur_nexact);
            if ((rc = launch_check())) break;
        }
        rc = step_epilogue(c, s + 1 == nsteps, need_pre);
    }
    c->write_v = wv;
    if (rc) return note_error(rc);
    return harvest_device_error(c);
}

int org_step_surrogate(org_ctx *c, fno_net *net, const float *v_given, int it0, int nsteps, int k_exact, float thr,
                       int32_t *exact_cells, float *residuals) {
    if (!c || c->gen || (!net && !v_given) || nsteps < 1 || k_exact < 1 || !(thr >= 0.0f)) return note_error(ORG_EINVAL);
    if (net) {
        fno_config fc;
        if (fno_get_config(net, &fc) || fc.M != ORG_MB || fc.batch_max < c->nc) return note_error(ORG_EINVAL);
    }
    TRY(need_prepared(c));
    TRY(sur_buffers(c, nsteps));
    TRY(step_prologue(c));
    void *old_stream = nullptr;
    if (net && (fno_get_stream(net, &old_stream) || fno_set_stream(net, (void *)c->stream)))
        return note_error(ORG_EINVAL);
    const float *vsrc = net ? c->sur_v : v_given;
    HIPCHK(hipMemsetAsync(c->sur_nexact, 0, sizeof(int) * nsteps, c->stream));
    int rc = ORG_OK;
    bool need_pre = true;
    for (int s = 0; s < nsteps && !rc; s++) {
        if (need_pre && (rc = L_pre(c))) break;
        float *res = c->sur_res + (size_t)s * c->nc;
        if ((it0 + s) % k_exact == 0) {
            hipLaunchKernelGGL(k_fill_use, dim3(1), dim3(256), 0, c->stream, c->sur_use, c->nc, 0,
                               c->sur_nexact + s, res);
            if ((rc = launch_check())) break;
        } else {
            if (net) {
                hipLaunchKernelGGL(k_crops, dim3(c->nc), dim3(1024), 0, c->stream, c->dp, c->phi, c->rho, c->xc,
                                   c->yc, c->sur_x);
                if ((rc = launch_check()) || (rc = fno_forward(net, c->sur_x, c->nc, c->sur_v))) break;
            }
            hipLaunchKernelGGL(c->judge == ORG_JUDGE_SPECTRAL ? k_vresid_spec : k_vresid, dim3(c->nc), dim3(1024), 0,
                               c->stream, c->dp, c->phi, c->rho, c->lum, c->ecm, c->xc, c->yc, c->spec, vsrc, thr,
                               res, c->sur_use, c->sur_nexact + s);
            if ((rc = launch_check())) break;
        }
        if ((rc = L_step_v(c, vsrc, c->sur_use))) break;
        rc = step_epilogue(c, s + 1 == nsteps, need_pre);
    }
    if (net) (void)fno_set_stream(net, old_stream);
    if (rc) return note_error(rc);
    if (exact_cells)
        HIPCHK(hipMemcpyAsync(exact_cells, c->sur_nexact, sizeof(int) * nsteps, hipMemcpyDeviceToHost, c->stream));
    if (residuals)
        HIPCHK(hipMemcpyAsync(residuals, c->sur_res, sizeof(float) * nsteps * c->nc, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return harvest_device_error(c);
}

int org_sync(org_ctx *c) {
    TRY(check_ctx(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    return ORG_OK;
}

int org_profile_enable(org_ctx *c, int on) {
    TRY(check_ctx(c));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->prof = on != 0;
    for (int k = 0; k < ORG_K_COUNT; k++) c->ev_used[k] = 0;
    return ORG_OK;
}

int org_kernel_time(org_ctx *c, int kid, double *ms, int64_t *launches) {
    TRY(check_ctx(c));
    if (kid < 0 || kid >= ORG_K_COUNT || !ms || !launches) return note_error(ORG_EINVAL);
    HIPCHK(hipStreamSynchronize(c->stream));
    double tot = 0.0;
    for (size_t k = 0; k < c->ev_used[kid]; k++) {
        float v = 0.0f;
        HIPCHK(hipEventElapsedTime(&v, c->ev[kid][k].first, c->ev[kid][k].second));
        tot += v;
    }
    *ms = tot;
    *launches = (int64_t)c->ev_used[kid];
    c->ev_used[kid] = 0;
    return ORG_OK;
}

/* ---- batches of organoids ------------------------------------------------ */
int org_batch_create(org_ctx *const *ctxs, int n, org_batch **out) {
    if (!ctxs || !out || n < 1 || n > ORG_BATCH_MAX) return note_error(ORG_EINVAL);
    *out = nullptr;
    for (int o = 0; o < n; o++) {
        if (!ctxs[o] || ctxs[o]->in_batch) return note_error(ORG_EINVAL);
        if (ctxs[o]->gen) {
            fprintf(stderr, "organoid: batches take FP32 contexts with zone_size %d only\n", ORG_ZS_MAX);
            return note_error(ORG_EINVAL);
        }
        for (int q = 0; q < o; q++)
            if (ctxs[q] == ctxs[o]) return note_error(ORG_EINVAL);
    }
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0)
        ncu = 256;
    /* role sizes: the per-cell roles have one workgroup per cell; the
     * others share the CUs in proportion to the grid areas.  The lumen/ECM
     * workgroups of k_step wait on each other, so all of them must be
     * resident at once: their total stays within the CU count. */
    double area = 0.0;
    long long cells = 0;
    for (int o = 0; o < n; o++) { area += (double)ctxs[o]->N * ctxs[o]->N; cells += ctxs[o]->nc; }
    const int lum_budget = std::max(n, std::min(ncu / 2, ORG_LUM_BLOCKS * n));
    const int spare = std::max(64, ncu - (int)std::min<long long>(cells, ncu));
    std::vector<int> nlum(n), role(n);
    int lum_tot = 0;
    for (int o = 0; o < n; o++) {
        const double w = (double)ctxs[o]->N * ctxs[o]->N / area;
        nlum[o] = std::max(1, std::min(ORG_LUM_BLOCKS, (int)(lum_budget * w)));
        role[o] = std::max(2, (int)std::lround(0.5 * spare * w));
        lum_tot += nlum[o];
    }
    if (lum_tot > ncu - 1) {
        fprintf(stderr, "organoid: a batch of %d organoids needs more lumen/ECM workgroups than CUs\n", n);
        return note_error(ORG_EINVAL);
    }
    org_batch *b = new org_batch();
    b->ctx.assign(ctxs, ctxs + n);
    int rc = ORG_OK;
    for (int o = 0; o < n && !rc; o++) {
        org_ctx *c = ctxs[o];
        if (hipStreamSynchronize(c->stream) != hipSuccess) rc = note_error(ORG_EHIP);
        c->nlum = nlum[o];
        c->npsi = role[o];
        c->nh = std::min(c->nparts_h, role[o]);
        c->nlp = std::min(c->nlp_cap, role[o]);
        c->next_ready = false;          /* partial live boxes are re-initialised for the new nlp */
        for (int k = 0; k < 2; k++) {    /* the context's own graphs hold the old role sizes */
            if (c->gexec[k]) { (void)hipGraphExecDestroy(c->gexec[k]); c->gexec[k] = nullptr; }
            if (c->graph[k]) { (void)hipGraphDestroy(c->graph[k]); c->graph[k] = nullptr; }
        }
        c->in_batch = true;
    }
    /* block maps: (role << 28) | (organoid << 20) | index */
    std::vector<int> maps[4];
    auto put = [](std::vector<int> &m, int role_, int o, int cnt) {
        for (int i = 0; i < cnt; i++) m.push_back((role_ << 28) | (o << 20) | i);
    };
    for (int o = 0; o < n; o++) {
        const org_ctx *c = ctxs[o];
        put(maps[BK_PRE], 0, o, c->nc); put(maps[BK_PRE], 1, o, c->npsi); put(maps[BK_PRE], 2, o, c->nh);
        put(maps[BK_STEP], 1, o, c->nlum);
        put(maps[BK_MID], 0, o, c->nc); put(maps[BK_MID], 1, o, c->npsi); put(maps[BK_MID], 2, o, c->nlp);
        put(maps[BK_POST], 0, o, c->nc); put(maps[BK_POST], 2, o, c->nlp);
    }
    for (int o = 0; o < n; o++) { put(maps[BK_STEP], 0, o, ctxs[o]->nc); put(maps[BK_STEP], 2, o, ctxs[o]->nc); }
    if (!rc && hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) rc = note_error(ORG_EHIP);
    if (!rc && hipMalloc(&b->ent, sizeof(BatchEnt) * n) != hipSuccess) rc = note_error(ORG_EHIP);
    for (int k = 0; k < 4 && !rc; k++) {
        b->grid[k] = (int)maps[k].size();
        if (hipMalloc(&b->bmap[k], sizeof(int) * maps[k].size()) != hipSuccess ||
            hipMemcpy(b->bmap[k], maps[k].data(), sizeof(int) * maps[k].size(), hipMemcpyHostToDevice) != hipSuccess)
            rc = note_error(ORG_EHIP);
    }
    if (!rc) rc = batch_upload_entries(b);
    if (rc) { org_batch_destroy(b); return rc; }
    *out = b;
    return ORG_OK;
}

int org_batch_destroy(org_batch *b) {
    if (!b) return ORG_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (org_ctx *c : b->ctx) c->in_batch = false;
    (void)hipFree(b->ent);
    for (int k = 0; k < 4; k++) (void)hipFree(b->bmap[k]);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return ORG_OK;
}

int org_batch_step(org_batch *b, int nsteps, int use_graph) {
    if (!b || nsteps < 0) return note_error(ORG_EINVAL);
    if (nsteps == 0) return ORG_OK;
    /* every organoid ready for the batched step (what org_step does first) */
    for (org_ctx *c : b->ctx) {
        if (!c->prepared) {
            fprintf(stderr, "organoid: org_prepare must run after the fields are uploaded\n");
            return note_error(ORG_ESTATE);
        }
        if (!c->next_ready) {
            TRY(L_centroid_next(c));
            TRY(L_live_init(c));
        }
        HIPCHK(hipStreamSynchronize(c->stream));
        c->next_ready = true;
    }
    TRY(batch_upload_entries(b));       /* options such as the velocity output may have changed */
    bool ring = false;
    for (org_ctx *c : b->ctx) ring = ring || c->ring_pending;
    if (ring) {                          /* first step after org_prepare: eager, then the ghost rings */
        TRY(enqueue_batch_steps(b, 1));
        HIPCHK(hipStreamSynchronize(b->stream));
        for (org_ctx *c : b->ctx) {
            TRY(L_ring_if_pending(c));
            HIPCHK(hipStreamSynchronize(c->stream));
        }
        nsteps--;
    }
    if (!use_graph) {
        TRY(enqueue_batch_steps(b, nsteps));
        return batch_harvest(b);
    }
    auto enq = [b](int n) { return enqueue_batch_steps(b, n); };
    const int full = nsteps / ORG_GRAPH_STEPS, rem = nsteps % ORG_GRAPH_STEPS;
    if (full) TRY(ensure_graph(b->stream, b->graph[0], b->gexec[0], b->graph_len[0], ORG_GRAPH_STEPS, enq));
    if (rem) TRY(ensure_graph(b->stream, b->graph[1], b->gexec[1], b->graph_len[1], rem, enq));
    for (int s = 0; s < full; s++) HIPCHK(hipGraphLaunch(b->gexec[0], b->stream));
    if (rem) HIPCHK(hipGraphLaunch(b->gexec[1], b->stream));
    return batch_harvest(b);
}

int org_batch_sync(org_batch *b) {
    if (!b) return note_error(ORG_EINVAL);
    HIPCHK(hipStreamSynchronize(b->stream));
    return ORG_OK;
}

/* ---- allocation-free h(lum), h(ECM) and sum of h(ECM) -------------------
 * The partial sums (one per 64 x 64 tile of the grid and one for the
 * boundary ring) and the last-workgroup counter of k_hlum live in a static
 * device array: each GPU holds its own copy from the library's load to the
 * process end.  Host code reaches a device's copy only through the address
 * hipGetSymbolAddress returns with that device current (the array's name is
 * a host-side stand-in, not a device address), looked up once per device. */
#define ORG_NOALLOC_MAX_N 4096
#define ORG_NOALLOC_PARTS ((ORG_NOALLOC_MAX_N / 64) * (ORG_NOALLOC_MAX_N / 64) + 1)
#define ORG_NOALLOC_BYTES (8 * ORG_NOALLOC_PARTS + 256)
static __device__ __align__(256) unsigned char scratch[ORG_NOALLOC_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};

/* This device's copy of `scratch` (looked up on the first call there). */
int scratch_address(void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return note_error(ORG_EINVAL);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "organoid: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return note_error(ORG_EHIP);
        }
        g_scratch_dev[dev] = p;
    }
    *out = g_scratch_dev[dev];
    return ORG_OK;
}

__global__ void k_zero_u32(unsigned *p) { *p = 0u; }
}  // namespace

int org_compute_hfunc_lum_ecm_noalloc(int N, double dx, const float *lum, const float *ecm, float *hlum, float *hecm,
                                      double *sum_ecm, void *stream) {
    if (!lum || !ecm || !hlum || !hecm || !sum_ecm || !(dx > 0.0)) return note_error(ORG_EINVAL);
    if (N < 8 || N > ORG_NOALLOC_MAX_N) {
        fprintf(stderr, "organoid: org_compute_hfunc_lum_ecm_noalloc: N = %d outside 8..%d (its scratch holds "
                        "the partial sums of at most %d x %d tiles)\n", N, ORG_NOALLOC_MAX_N, ORG_NOALLOC_MAX_N / 64,
                ORG_NOALLOC_MAX_N / 64);
        return note_error(ORG_EINVAL);
    }
    void *base = nullptr;
    TRY(scratch_address(&base));
    const int t = (N + 63) / 64, nparts = t * t + 1;
    double *part = (double *)base;
    unsigned *counter = (unsigned *)((unsigned char *)base + 8 * ORG_NOALLOC_PARTS);
    DevP d{};
    d.N = N;
    d.dx = (float)dx;
    hipStream_t st = (hipStream_t)stream;
    /* the counter must start at 0: a one-thread kernel rather than a memset,
     * so the call can be captured into a graph */
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, st, counter);
    TRY(launch_check());
    hipLaunchKernelGGL(k_hlum, dim3(nparts), dim3(1024), 0, st, d, lum, ecm, hlum, hecm, part, counter, sum_ecm);
    return launch_check();
}

int org_debug_fft2d(float *data, int sign) {
    if (!data || (sign != 1 && sign != -1)) return note_error(ORG_EINVAL);
    float2 *d = nullptr;
    size_t bytes = sizeof(float2) * ORG_MB * ORG_MB;
    HIPCHK(hipMalloc(&d, bytes));
    int rc = ORG_OK;
    if (hipMemcpy(d, data, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = note_error(ORG_EHIP);
    if (!rc) {
        hipLaunchKernelGGL(k_fft_test, dim3(1), dim3(1024), 0, 0, d, sign);
        rc = launch_check();
    }
    if (!rc && hipMemcpy(data, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = note_error(ORG_EHIP);
    (void)hipFree(d);
    return rc;
}

int org_debug_fft2d_generic(int M, double *data, int sign, int f64, int fuse, int lines_per_wg) {
    return gen_rc(gen::debug_fft2d(M, data, sign, f64, fuse, lines_per_wg));
}

int org_get_fft_schedule(org_ctx *c, int32_t *fuse, int32_t *lines_per_wg) {
    TRY(check_ctx(c));
    if (!fuse || !lines_per_wg) return note_error(ORG_EINVAL);
    if (!c->gen) {                 /* the LDS transform: 16 x 8, one 128^2 plane per workgroup */
        *fuse = 0;
        *lines_per_wg = 0;
        return ORG_OK;
    }
    return gen_rc(c->gen->get_fft_schedule(fuse, lines_per_wg));
}

int org_set_fft_schedule(org_ctx *c, int fuse, int lines_per_wg) {
    TRY(check_ctx(c));
    if (!c->gen) return note_error(ORG_EINVAL);
    return gen_rc(c->gen->set_fft_schedule(fuse, lines_per_wg));
}

/* ---- double-precision host buffers ----------------------------------------
 * FP64 contexts move doubles unchanged; FP32 contexts round to float on the
 * way in and widen on the way out. */
static size_t field_count(org_ctx *c, int field) {
    switch (field) {
    case ORG_F_PHI: case ORG_F_RHO: case ORG_F_HFUNC: return c->P;
    case ORG_F_LUM: case ORG_F_ECM: case ORG_F_PSI: case ORG_F_HLUM: case ORG_F_HECM:
    case ORG_F_LUMN: case ORG_F_ECMN: return c->NN;
    default: return 0;
    }
}

int org_set_field_f64(org_ctx *c, int field, const double *host) {
    TRY(check_ctx(c));
    if (!host) return note_error(ORG_EINVAL);
    if (c->gen) {
        TRY(gen_rc(c->gen->set_field(field, host, 1)));
        unprepare_on(c, field);
        return ORG_OK;
    }
    const size_t n = field_count(c, field);
    if (!n) return note_error(ORG_EINVAL);
    std::vector<float> t(n);
    for (size_t k = 0; k < n; k++) t[k] = (float)host[k];
    return org_set_field(c, field, t.data());
}

int org_get_field_f64(org_ctx *c, int field, double *host) {
    TRY(check_ctx(c));
    if (!host) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_field(field, host, 1));
    const size_t n = field_count(c, field);
    if (!n) return note_error(ORG_EINVAL);
    std::vector<float> t(n);
    TRY(org_get_field(c, field, t.data()));
    for (size_t k = 0; k < n; k++) host[k] = t[k];
    return ORG_OK;
}

int org_get_window_f64(org_ctx *c, int which, double *host) {
    TRY(check_ctx(c));
    if (!host) return note_error(ORG_EINVAL);
    if (c->gen) return gen_rc(c->gen->get_window(which, host, 1));
    if (which < 0 || which >= ORG_W_COUNT) return note_error(ORG_EINVAL);
    const size_t n = (size_t)c->nc * (which == ORG_W_PHIN ? (size_t)c->W * c->W : (size_t)c->M * c->M);
    std::vector<float> t(n);
    TRY(org_get_window(c, which, t.data()));
    for (size_t k = 0; k < n; k++) host[k] = t[k];
    return ORG_OK;
}

}  /* extern "C" */
