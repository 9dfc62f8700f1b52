// th_internal.hpp -- host-side internals shared by the translation units of libtarang_hip.so.
#pragma once
#include <hip/hip_runtime.h>
#include <map>
#include <atomic>
#include <mutex>
#include <string>
#include <vector>
#include "th_common.hpp"

struct th_force;

struct th_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string tune_path;                     // tune cache file ("" = in memory only)
    std::mutex tune_mu;
    std::map<std::string, std::vector<int>> tune;   // launch-shape choices by workload key
    std::atomic<int> nplans{0};
};

namespace th {
// Tune cache: one line per workload, "<key> <v0> <v1> ...", read when the context is created and
// rewritten whenever the autotuner adds an entry.
bool tune_lookup(th_ctx* c, const std::string& key, std::vector<int>& v);
void tune_store(th_ctx* c, const std::string& key, const std::vector<int>& v);
}  // namespace th

namespace th {

// status codes (mirrored in include/tarang_hip.h)
enum { OK = 0, E_ARG = -1, E_HIP = -2, E_UNSUPPORTED = -3, E_STATE = -4, E_NOMEM = -5, E_IO = -6 };

void set_error(const char* fmt, ...);

// propagate a nonzero status code
#define TRY(x)             \
    do {                   \
        int r_ = (x);      \
        if (r_) return r_; \
    } while (0)

#define TH_CHECK_HIP(expr)                                                                 \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) {                                                            \
            th::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_)); \
            return th::E_HIP;                                                              \
        }                                                                                  \
    } while (0)

// Geometry of one box (reference layout (N1, N2, N3/2+1) complex, C-order).
struct Geo {
    int basis = 0;          // 0 FOUR, 1 SCFT
    int N1 = 0, N2 = 0, N3 = 0, H = 0;
    double kf[3] = {1, 1, 1};
    int dealias = 1;
    size_t ncplx() const { return (size_t)N1 * N2 * H; }
};

// Twiddle tables W_L^m = exp(-2 pi i m / L), m < L, built from FP64 libm on the host.
struct TwiddleCache {
    std::map<int, void*> f32, f64;
    const void* get(int L, int prec /*0 fp64, 1 fp32*/, hipStream_t s);
    void release();
};

// ---------------- generic (plain) kernels: th_generic.hip ----------------
template <typename R> int fft_strided(cplx<R>* data, long sO, long sL, int O, int L, int M, int sign, R scale,
                                      const cplx<R>* tw, hipStream_t st);
template <typename R> int fft_z_r2c(cplx<R>* data, long nlines, int N3, R scale, const cplx<R>* twL,
                                    const cplx<R>* twN, hipStream_t st);
template <typename R> int fft_z_c2r(cplx<R>* data, long nlines, int N3, const cplx<R>* twL, const cplx<R>* twN,
                                    hipStream_t st);
template <typename R> int r2r_x(cplx<R>* data, const Geo& g, int kind, cplx<R>* work, const cplx<R>* tw2N,
                                const cplx<R>* tw4N, hipStream_t st);
template <typename R> int shift_x(cplx<R>* data, const Geo& g, int dir, hipStream_t st);

// elementwise spectral ops on reference-layout arrays
template <typename R> int k_deriv(const Geo& g, int axis, int parity, const cplx<R>* A, cplx<R>* B, hipStream_t st);
template <typename R> int k_add_deriv(const Geo& g, int axis, int parity, const cplx<R>* A, cplx<R>* B,
                                      hipStream_t st);       // B += D(A)
template <typename R> int k_dealias(const Geo& g, cplx<R>* A, hipStream_t st);
template <typename R> int k_real_mult(size_t n, const cplx<R>* A, const cplx<R>* B, cplx<R>* C, hipStream_t st);
template <typename R> int k_elsasser(size_t n, cplx<R>* U, cplx<R>* B, int back, hipStream_t st);
template <typename R> int k_divide_ksqr(const Geo& g, cplx<R>* A, hipStream_t st);
template <typename R> int k_exp_ksqr(const Geo& g, cplx<R>* A, double factor, hipStream_t st);
template <typename R> int k_axpy(size_t n, cplx<R>* Y, double a, const cplx<R>* X, hipStream_t st);   // Y = Y + a X
template <typename R> int k_axpby(size_t n, cplx<R>* Y, double a, const cplx<R>* X, double b, hipStream_t st); // Y = a X + b Y
template <typename R> int k_scale(size_t n, cplx<R>* Y, double a, hipStream_t st);
template <typename R> int k_fill_shell(const Geo& g, const cplx<R>* A, cplx<R>* B, double inner, double outer,
                                       hipStream_t st);
template <typename R> int k_zero_plane_x0(const Geo& g, cplx<R>* A, hipStream_t st);
template <typename R> int k_zero_padding(const Geo& g, cplx<R>* A, hipStream_t st);

// ---------------- reductions: th_reduce.hip ----------------
struct RedScratch {
    double* d = nullptr;
    size_t cap = 0;
    double* h = nullptr;
    size_t hcap = 0;
    int ensure(size_t n_dev, size_t n_host);
    void release();
};
// binned sum over modes: out[index] += w(mode) * f(mode); modes with index outside [0, nbins) skipped
// mode selects the weight/binning rule (spectrum, shell_mult_all, shell_mult_single ...)
template <typename R> int reduce_shell(const Geo& g, const cplx<R>* A, const cplx<R>* B, double n, int nbins,
                                       double* Sk, double* cnt, RedScratch& rs, hipStream_t st, bool rescale = true);
// Rings of the ring-to-ring (shells x polar sectors) and cylinder-ring-to-ring (K_perp shells x K_pll slabs)
// transfers (compute_ring_tr.cc, compute_cylin_ring_tr.cc): radii r[0..nr], sector angles / slab edges e[0..nsec]
constexpr int RING_MAX = 32;
struct RingP {
    int cyl, axis, nr, nsec, kin, kpm;
    double r[RING_MAX + 1], e[RING_MAX + 1];
};
int ring_params(const Geo& g, int cyl, int axis, const double* radii, int nr, const double* edges, int nsec, RingP* rp);
// G = F on giver ring (gs, gq) (Fill_array_ring / Fill_array_cylinder_ring universal_ET.cc:70-127), else 0
template <typename R> int fill_ring(const Geo& g, const cplx<R>* F, cplx<R>* G, const RingP& rp, int gs, int gq,
                                    hipStream_t st);
// out[(nr+1)(nsec+1)] = Ring_mult_all / Cyl_ring_mult_all of A, B (1 or 3 components; A[1] == nullptr: one)
template <typename R> int reduce_ring_mult(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3],
                                           const RingP& rp, double* out, RedScratch& rs, hipStream_t st);
// reduce_shell with no allocation: partials and results in a file-scope device array (th_reduce.hip)
template <typename R> int reduce_shell_noalloc(const Geo& g, const cplx<R>* A, const cplx<R>* B, double n, int nbins,
                                               double* Sk, double* cnt, hipStream_t st);
template <typename R> int reduce_energy(const Geo& g, const cplx<R>* A, const cplx<R>* B, double* out,
                                        RedScratch& rs, hipStream_t st);
template <typename R> int reduce_Sn(const Geo& g, const cplx<R>* A, double n, double* out, RedScratch& rs,
                                    hipStream_t st);
template <typename R> int reduce_shell_mult_single(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3],
                                                   double inner, double outer, double* out, RedScratch& rs,
                                                   hipStream_t st);
template <typename R> int reduce_shell_mult_all(const Geo& g, const cplx<R>* const A[3], const cplx<R>* const B[3],
                                                const double* radii_dev, int nr, double* out, RedScratch& rs,
                                                hipStream_t st);
template <typename R> int reduce_helicity(const Geo& g, const cplx<R>* const V[3], double* out, RedScratch& rs, hipStream_t st);
template <typename R> int reduce_entropy(const Geo& g, const cplx<R>* const V[3], double* out, RedScratch& rs, hipStream_t st);
template <typename R> int reduce_entropy_scalar(const Geo& g, const cplx<R>* A, double* out, RedScratch& rs, hipStream_t st);
template <typename R> int reduce_shell_helicity(const Geo& g, const cplx<R>* const V[3], int nbins, double* H1, double* H2,
                                                double* H3, RedScratch& rs, hipStream_t st);

// ---------------- energy / helicity forcing: th_force.hip ----------------
// The band inner < |K| <= outer of force_field_proc 1 / 2, its per-mode level and the bounding box of
// wavevectors one launch covers.
struct ForceBand {
    int N1, N2, N3, H;
    double f1, f2, f3;
    int proc;                 // 1 energy/helicity supply, 2 constant energy/helicity
    double inner, outer;
    double per_mode;          // energy_supply / nf (proc 1) or energy_level / nf (proc 2)
    double p4;                // epsh_by_k_epse (proc 1) or h_by_k_E (proc 2)
    long nf;                  // modes in inner <= |K| < outer, conjugates included
    int dealias_force;        // Is_alias_array: the band reaches past N/3 on some axis
    int fullx, fully;         // the box spans every x (y) index
    int kxm, kym;             // box origin: kx = tx - kxm, ky = ty - kym
    int nx, ny, nz;           // box extents
    double emin;              // a mode is forced when its modal energy exceeds this (MYEPS; 0 for W, proc 2)
    int k0;                   // proc 3 / 4 (Taylor-Green, ABC): the forced wavenumber; per_mode holds the amplitude
    double abc[3];            // proc 4: A, B, C
};
int force_setup(const Geo& g, const th_force* f, ForceBand& b);
// the bands of U (p3, p4) and W (p5, p6) of Compute_force(IncVF& W) (compute_force_ek_hk_supply.cc:294-420,
// compute_force_const_ek_hk.cc:296-437)
int force_setup_mhd(const Geo& g, const th_force* f, ForceBand& bu, ForceBand& bw);
// F = the force of V (zero outside the band)
template <typename R> int force_write(const ForceBand& b, const cplx<R>* const V[3], cplx<R>* const F[3], size_t n,
                                      hipStream_t st);
// nl[c] -= F[c] over the band (Add_force after Compute_force)
template <typename R> int force_subtract(const ForceBand& b, const cplx<R>* const V[3], cplx<R>* const nl[3],
                                         hipStream_t st);

// anisotropic spectra: mode 0 ring (rows Max_radius_inside+1), 1 cylinder (rows cylinder_shell_ek_size);
// S1, S2 host arrays of rows x (nbin+1)
template <typename R> int reduce_aniso(const Geo& g, const cplx<R>* const V[3], int mode, int nbin, int n, int axis,
                                       double* S1, double* S2, RedScratch& rs, hipStream_t st);
int cylinder_shell_ek_size(const Geo& g, int axis);

#ifdef __HIP__
// out[b] = sum over k < nblk of partial[k * stride + b], b < nval: one workgroup per value, each thread
// a compensated (Neumaier) sum over a strided subset of the k, then the (sum, compensation) pairs
// combined in a fixed tree, so the result does not depend on scheduling.
constexpr int FIN_T = 256;
static __global__ __launch_bounds__(FIN_T) void k_finalize_bins(const double* __restrict__ partial, int nblk, long stride,
                                                                double* __restrict__ out)
{
    __shared__ double ss[FIN_T], cc[FIN_T];
    const int b = blockIdx.x, t = threadIdx.x;
    double s = 0.0, c = 0.0;
    for (int k = t; k < nblk; k += FIN_T) {
        const double v = partial[(long)k * stride + b];
        const double u = s + v;
        c += (fabs(s) >= fabs(v)) ? ((s - u) + v) : ((v - u) + s);
        s = u;
    }
    ss[t] = s;
    cc[t] = c;
    __syncthreads();
    for (int w = FIN_T / 2; w > 0; w >>= 1) {
        if (t < w) {
            const double s1 = ss[t], s2 = ss[t + w], u = s1 + s2;
            const double e = (fabs(s1) >= fabs(s2)) ? ((s1 - u) + s2) : ((s2 - u) + s1);
            ss[t] = u;
            cc[t] = cc[t] + cc[t + w] + e;
        }
        __syncthreads();
    }
    if (t == 0) out[b] = ss[0] + cc[0];
}
static inline void launch_finalize(const double* partial, int nblk, long stride, int nval, double* out, hipStream_t st)
{
    hipLaunchKernelGGL(k_finalize_bins, dim3((unsigned)nval), dim3(FIN_T), 0, st, partial, nblk, stride, out);
}
#endif

// host-side binning helpers (four_inline.h, scft_inline.h)
int min_radius_outside(const Geo& g);
int max_radius_inside(const Geo& g);
double approx_modes_in_shell(const Geo& g, int radius);

}  // namespace th
