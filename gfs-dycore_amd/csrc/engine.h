// Engine<T>: the device-resident transform and dycore machinery behind one
// gfsdc context, at compute precision T (double = FP64 path, float = FP32 path).
#pragma once
#include <memory>

#include "gfsdc_internal.h"

namespace gfsdc {

// Grid-space fields held by the dycore after getdyntend (grid_data.f90 names).
enum GridField : int {
    GF_UG = 0, GF_VG, GF_VRT, GF_DIV, GF_TV, GF_DLNPDT, GF_ETADOT, GF_PSG, GF_DLNPSDT, GF_PRS,
    GF_DLNPSDX = 14, GF_DLNPSDY = 15, GF_TRACER = 16, GF_PRECIP = 17, GF_PWAT = 18
};

struct DycoreBuffers;  // defined in dycore.hip

template <typename T> struct Engine {
    gfsdc_cfg cfg{};
    std::vector<double> ak, bk;
    hipStream_t stream = nullptr;
    Plan<T> plan;
    Work<T> work;                 // per-field API workspace
    // staging for caller buffers
    void *stage[4] = {nullptr, nullptr, nullptr, nullptr};
    size_t stage_sz[4] = {0, 0, 0, 0};
    SynthDesc *d_sdesc = nullptr;
    AnalDesc *d_adesc = nullptr;
    unsigned char *d_rflag = nullptr;
    int desc_cap = 0;
    std::vector<SynthDesc> h_sdesc;
    std::vector<AnalDesc> h_adesc;
    std::vector<unsigned char> h_rflag;
    DycoreBuffers *dyn = nullptr;

    int init(const gfsdc_cfg &c);
    void destroy();
    int ensure_stage(int i, size_t bytes);
    int upload_descs(const std::vector<SynthDesc> &sd, const std::vector<AnalDesc> &ad,
                     const std::vector<unsigned char> &rf);
    // caller buffer (n real values of dtype at loc) -> device T buffer
    int input(const void *p, size_t n, int dtype, int loc, int slot, const T **out);
    // device T buffer -> caller buffer
    int output(const T *d, void *p, size_t n, int dtype, int loc, int slot);

    int spectogrd(const void *spec, void *grid, int nf, int dtype, int loc);
    int spectogrd_noalloc(const void *spec, void *grid, int nf);
    int bench_transforms(int ns, int nv, int reps, double *ms);
    int bench_copy(long nbytes, int reps, double *ms);
    int grdtospec(const void *grid, void *spec, int nf, int dtype, int loc);
    int sphtor_to_spat(const void *s, const void *t, void *u, void *v, int nf, double sscale, int sinv,
                       double tscale, int tinv, int dtype, int loc);
    int spat_to_sphtor(const void *u, const void *v, void *s, void *t, int nf, double rsphere, int inv,
                       int dtype, int loc);

    // dycore (dycore.hip)
    int dyn_init();
    void dyn_destroy();
    int set_state(const void *vrt, const void *div, const void *tv, const void *trac, const void *lnps,
                  int dtype, int loc);
    int get_state(void *vrt, void *div, void *tv, void *trac, void *lnps, int dtype, int loc);
    int set_topo(const void *topo, int dtype, int loc);
    int getdyntend(int rkstage, int inverse_only, void *dvrt, void *ddiv, void *dtv, void *dtrac,
                   void *dlnps, int dtype, int loc);
    int advance(int nsteps, int use_graph);
    int hs_phytend(void *dvrt, void *ddiv, void *dtv, int dtype, int loc);
    int hs_drag_levels(int *n);
    int simple_phytend(void *dvrt, void *ddiv, void *dtv, void *dtrac, int dtype, int loc);
    int pack_tphys(void *dvrt, void *ddiv, void *dtv, void *dtrac, int dtype, int loc);
    int pattern_spinup();
    int autotune(double *ms, int *best);
    int get_pattern(void *out, int dtype, int loc);
    // learned column physics
    int mlp_init(int hidden, int nlayers, int maxbatch, long capacity, unsigned long long seed);
    int mlp_collect(int nsteps, int stride, long *ncollected);
    int mlp_normalize();
    int mlp_train(int nsteps, int batch, double lr, double wd, long ntrain, double *losses);
    int mlp_eval(long first, long count, double *err, double *base);
    int mlp_enable(int on, int judge_stride, double threshold);
    int mlp_judge_stats(double *out);
    long mlp_nparams();
    int mlp_params(float *get, const float *set);
    int mlp_norm(float *get, const float *set);
    int mlp_forward_host(const float *x, long n, float *y);
    int mlp_grad_host(const float *x, const float *t, int n, float *grad, double *loss);
    int mlp_adamw_host(const float *grad, double lr, double wd);
    int get_grid(int which, void *out, int dtype, int loc);
    int get_semimp(double *amhyb, double *bmhyb, double *svhyb, double *tor_hyb, double *d_hyb_m);
    int get_damp(double *disspec, double *diff_prof, double *dmp_prof);
    int step_diag(double *out);
    int set_pdryini(const double *p);
    int get_pdryini(double *p);
    int init_rest_state(double t0, double ps0, double vort_rms, int nmin, int nmax, unsigned long long seed);
    int set_time(double t);
    int get_time(double *t);
    int iau_set(int nfiles, const double *fhrs, double delthrs, const void *vrt, const void *div, const void *tv,
                const void *trac, const void *lnps, int dtype, int loc);
    int dfi(int ntdfi, int use_graph);
    int stoch_init(const gfsdc_stoch_cfg *sc);
    int get_stoch(int which, void *out, int dtype, int loc);
    size_t checkpoint_size();
    int checkpoint_save(void *buf, size_t nbytes, int loc);
    int checkpoint_load(const void *buf, size_t nbytes, int loc);
    int profile_step(char names[][48], double *ms, int nmax);
};

}  // namespace gfsdc

struct gfsdc_ctx {
    int precision = GFSDC_FP64;
    gfsdc::Engine<float> *ef = nullptr;
    gfsdc::Engine<double> *ed = nullptr;
};
