// Frequency-domain part of the batched history convolution (gfx950).
//
// The convolution of step k sums, for every pair (s, f), the history blocks
// B_j of raw lag j times the reference's 1-2-1 combination of states,
//   F_x = g_{x-1} + 2 g_x + g_{x+1}   (g_m = 0 outside 1 <= m < nstore),
// as sum_j B_j F_{k-j} (DBEM1/multidomain.cpp:3201-3225,3849-3866).  The raw
// lags j <= 16 stay with the lag-by-lag kernels (8-step tiles, batched.hip);
// the rest is split by raw lag into levels l = 0, 1, ... with sub-filter
// length L_l = 16 * 2^l:
//   level l < last: lags L_l + 1 .. 2 L_l (one sub-filter),
//   last level:     lags L + 1 .. maxn (sub-filters q = 1, 2, ...),
// sub-filter q of a level holding the lags L q + 1 .. L q + L.  A level runs
// once per L-step tile (steps K0 + i, i < L, K0 = 2 + L t); every state its
// lags reach there was solved before the tile (k - j + 1 <= K0 - 1 for
// j >= L + 1).  Per level this is the uniformly partitioned overlap-save
// convolution: with 2L-point spectra H_q of the zero-padded sub-filters and
// W_beta of the state windows (F at L beta - 2L + 1 .. L beta), tile t gets
// samples L .. 2L - 1 of the inverse transform of Y = sum_q H_q W_{t+1-q}.
// Per (pair, sub-filter, tile) that is L + 1 spectral bins x 18 complex
// products instead of L lags x L steps x 18 block products, so the longer
// levels take the long tail of every band at a fraction of the multiply-adds
// (the partition is non-uniform: short sub-filters where the lags are short,
// so that no level needs a state of its own tile).  All in FP64; the
// transforms round at 1e-16.  The L + 1 real-input bins are held as L
// complex slots: slot 0 packs the real bins 0 and L as (re, im).
//
// Kernels per level: the per row-group list of (field, sub-filter) items and
// the sub-filter spectra of every item (once per solve); the state-window
// spectra of every history column and load case (once per tile); the
// spectral products, where a workgroup of eight wavefronts takes one slot,
// 64 source rows of one domain (eight per wavefront) and 64 load cases, so a
// window spectrum fetched for an item serves every row of the group; and the
// inverse transforms of the summed spectra (once per tile; level 0's adds the
// other levels' samples of its 16 steps).
#include "batched.hpp"

namespace tdbem {

namespace {

// (cos, sin) of 2 pi k / 256, k = 0 .. 255 (first-quadrant values, signs by
// symmetry); a 2L-point transform takes entry (k mod 2L) * (128 / L)
__constant__ double2 c_tw[256] = {
    {1.0, 0.0}, {0.999698818696204220116, 0.024541228522912288032},
    {0.998795456205172392715, 0.049067674327418014255}, {0.997290456678690216136, 0.073564563599667423529},
    {0.995184726672196886245, 0.098017140329560601994}, {0.992479534598709998157, 0.122410675199216198499},
    {0.989176509964780973452, 0.146730474455361751659}, {0.985277642388941244774, 0.170961888760301226364},
    {0.980785280403230449126, 0.195090322016128267848}, {0.975702130038528544460, 0.219101240156869797228},
    {0.970031253194543992604, 0.242980179903263889948}, {0.963776065795439866686, 0.266712757474898386325},
    {0.956940335732208864936, 0.290284677254462367636}, {0.949528180593036667196, 0.313681740398891476656},
    {0.941544065183020778413, 0.336889853392220050689}, {0.932992798834738887712, 0.359895036534988148775},
    {0.923879532511286756128, 0.382683432365089771728}, {0.914209755703530654635, 0.405241314004989870908},
    {0.903989293123443331586, 0.427555093430282094321}, {0.893224301195515320342, 0.449611329654606600046},
    {0.881921264348355029713, 0.471396736825997648556}, {0.870086991108711418652, 0.492898192229784036873},
    {0.857728610000272069902, 0.514102744193221726594}, {0.844853565249707073260, 0.534997619887097210663},
    {0.831469612302545237079, 0.555570233019602224743}, {0.817584813151583696505, 0.575808191417845300746},
    {0.803207531480644909807, 0.595699304492433343467}, {0.788346427626606262009, 0.615231590580626845485},
    {0.773010453362736960811, 0.634393284163645498215}, {0.757208846506484547575, 0.653172842953776764084},
    {0.740951125354959091176, 0.671558954847018400625}, {0.724247082951466920941, 0.689540544737066924617},
    {0.707106781186547524401, 0.707106781186547524401}, {0.689540544737066924617, 0.724247082951466920941},
    {0.671558954847018400625, 0.740951125354959091176}, {0.653172842953776764084, 0.757208846506484547575},
    {0.634393284163645498215, 0.773010453362736960811}, {0.615231590580626845485, 0.788346427626606262009},
    {0.595699304492433343467, 0.803207531480644909807}, {0.575808191417845300746, 0.817584813151583696505},
    {0.555570233019602224743, 0.831469612302545237079}, {0.534997619887097210663, 0.844853565249707073260},
    {0.514102744193221726594, 0.857728610000272069902}, {0.492898192229784036873, 0.870086991108711418652},
    {0.471396736825997648556, 0.881921264348355029713}, {0.449611329654606600046, 0.893224301195515320342},
    {0.427555093430282094321, 0.903989293123443331586}, {0.405241314004989870908, 0.914209755703530654635},
    {0.382683432365089771728, 0.923879532511286756128}, {0.359895036534988148775, 0.932992798834738887712},
    {0.336889853392220050689, 0.941544065183020778413}, {0.313681740398891476656, 0.949528180593036667196},
    {0.290284677254462367636, 0.956940335732208864936}, {0.266712757474898386325, 0.963776065795439866686},
    {0.242980179903263889948, 0.970031253194543992604}, {0.219101240156869797228, 0.975702130038528544460},
    {0.195090322016128267848, 0.980785280403230449126}, {0.170961888760301226364, 0.985277642388941244774},
    {0.146730474455361751659, 0.989176509964780973452}, {0.122410675199216198499, 0.992479534598709998157},
    {0.098017140329560601994, 0.995184726672196886245}, {0.073564563599667423529, 0.997290456678690216136},
    {0.049067674327418014255, 0.998795456205172392715}, {0.024541228522912288032, 0.999698818696204220116},
    {0.0, 1.0}, {-0.024541228522912288032, 0.999698818696204220116},
    {-0.049067674327418014255, 0.998795456205172392715}, {-0.073564563599667423529, 0.997290456678690216136},
    {-0.098017140329560601994, 0.995184726672196886245}, {-0.122410675199216198499, 0.992479534598709998157},
    {-0.146730474455361751659, 0.989176509964780973452}, {-0.170961888760301226364, 0.985277642388941244774},
    {-0.195090322016128267848, 0.980785280403230449126}, {-0.219101240156869797228, 0.975702130038528544460},
    {-0.242980179903263889948, 0.970031253194543992604}, {-0.266712757474898386325, 0.963776065795439866686},
    {-0.290284677254462367636, 0.956940335732208864936}, {-0.313681740398891476656, 0.949528180593036667196},
    {-0.336889853392220050689, 0.941544065183020778413}, {-0.359895036534988148775, 0.932992798834738887712},
    {-0.382683432365089771728, 0.923879532511286756128}, {-0.405241314004989870908, 0.914209755703530654635},
    {-0.427555093430282094321, 0.903989293123443331586}, {-0.449611329654606600046, 0.893224301195515320342},
    {-0.471396736825997648556, 0.881921264348355029713}, {-0.492898192229784036873, 0.870086991108711418652},
    {-0.514102744193221726594, 0.857728610000272069902}, {-0.534997619887097210663, 0.844853565249707073260},
    {-0.555570233019602224743, 0.831469612302545237079}, {-0.575808191417845300746, 0.817584813151583696505},
    {-0.595699304492433343467, 0.803207531480644909807}, {-0.615231590580626845485, 0.788346427626606262009},
    {-0.634393284163645498215, 0.773010453362736960811}, {-0.653172842953776764084, 0.757208846506484547575},
    {-0.671558954847018400625, 0.740951125354959091176}, {-0.689540544737066924617, 0.724247082951466920941},
    {-0.707106781186547524401, 0.707106781186547524401}, {-0.724247082951466920941, 0.689540544737066924617},
    {-0.740951125354959091176, 0.671558954847018400625}, {-0.757208846506484547575, 0.653172842953776764084},
    {-0.773010453362736960811, 0.634393284163645498215}, {-0.788346427626606262009, 0.615231590580626845485},
    {-0.803207531480644909807, 0.595699304492433343467}, {-0.817584813151583696505, 0.575808191417845300746},
    {-0.831469612302545237079, 0.555570233019602224743}, {-0.844853565249707073260, 0.534997619887097210663},
    {-0.857728610000272069902, 0.514102744193221726594}, {-0.870086991108711418652, 0.492898192229784036873},
    {-0.881921264348355029713, 0.471396736825997648556}, {-0.893224301195515320342, 0.449611329654606600046},
    {-0.903989293123443331586, 0.427555093430282094321}, {-0.914209755703530654635, 0.405241314004989870908},
    {-0.923879532511286756128, 0.382683432365089771728}, {-0.932992798834738887712, 0.359895036534988148775},
    {-0.941544065183020778413, 0.336889853392220050689}, {-0.949528180593036667196, 0.313681740398891476656},
    {-0.956940335732208864936, 0.290284677254462367636}, {-0.963776065795439866686, 0.266712757474898386325},
    {-0.970031253194543992604, 0.242980179903263889948}, {-0.975702130038528544460, 0.219101240156869797228},
    {-0.980785280403230449126, 0.195090322016128267848}, {-0.985277642388941244774, 0.170961888760301226364},
    {-0.989176509964780973452, 0.146730474455361751659}, {-0.992479534598709998157, 0.122410675199216198499},
    {-0.995184726672196886245, 0.098017140329560601994}, {-0.997290456678690216136, 0.073564563599667423529},
    {-0.998795456205172392715, 0.049067674327418014255}, {-0.999698818696204220116, 0.024541228522912288032},
    {-1.0, 0.0}, {-0.999698818696204220116, -0.024541228522912288032},
    {-0.998795456205172392715, -0.049067674327418014255}, {-0.997290456678690216136, -0.073564563599667423529},
    {-0.995184726672196886245, -0.098017140329560601994}, {-0.992479534598709998157, -0.122410675199216198499},
    {-0.989176509964780973452, -0.146730474455361751659}, {-0.985277642388941244774, -0.170961888760301226364},
    {-0.980785280403230449126, -0.195090322016128267848}, {-0.975702130038528544460, -0.219101240156869797228},
    {-0.970031253194543992604, -0.242980179903263889948}, {-0.963776065795439866686, -0.266712757474898386325},
    {-0.956940335732208864936, -0.290284677254462367636}, {-0.949528180593036667196, -0.313681740398891476656},
    {-0.941544065183020778413, -0.336889853392220050689}, {-0.932992798834738887712, -0.359895036534988148775},
    {-0.923879532511286756128, -0.382683432365089771728}, {-0.914209755703530654635, -0.405241314004989870908},
    {-0.903989293123443331586, -0.427555093430282094321}, {-0.893224301195515320342, -0.449611329654606600046},
    {-0.881921264348355029713, -0.471396736825997648556}, {-0.870086991108711418652, -0.492898192229784036873},
    {-0.857728610000272069902, -0.514102744193221726594}, {-0.844853565249707073260, -0.534997619887097210663},
    {-0.831469612302545237079, -0.555570233019602224743}, {-0.817584813151583696505, -0.575808191417845300746},
    {-0.803207531480644909807, -0.595699304492433343467}, {-0.788346427626606262009, -0.615231590580626845485},
    {-0.773010453362736960811, -0.634393284163645498215}, {-0.757208846506484547575, -0.653172842953776764084},
    {-0.740951125354959091176, -0.671558954847018400625}, {-0.724247082951466920941, -0.689540544737066924617},
    {-0.707106781186547524401, -0.707106781186547524401}, {-0.689540544737066924617, -0.724247082951466920941},
    {-0.671558954847018400625, -0.740951125354959091176}, {-0.653172842953776764084, -0.757208846506484547575},
    {-0.634393284163645498215, -0.773010453362736960811}, {-0.615231590580626845485, -0.788346427626606262009},
    {-0.595699304492433343467, -0.803207531480644909807}, {-0.575808191417845300746, -0.817584813151583696505},
    {-0.555570233019602224743, -0.831469612302545237079}, {-0.534997619887097210663, -0.844853565249707073260},
    {-0.514102744193221726594, -0.857728610000272069902}, {-0.492898192229784036873, -0.870086991108711418652},
    {-0.471396736825997648556, -0.881921264348355029713}, {-0.449611329654606600046, -0.893224301195515320342},
    {-0.427555093430282094321, -0.903989293123443331586}, {-0.405241314004989870908, -0.914209755703530654635},
    {-0.382683432365089771728, -0.923879532511286756128}, {-0.359895036534988148775, -0.932992798834738887712},
    {-0.336889853392220050689, -0.941544065183020778413}, {-0.313681740398891476656, -0.949528180593036667196},
    {-0.290284677254462367636, -0.956940335732208864936}, {-0.266712757474898386325, -0.963776065795439866686},
    {-0.242980179903263889948, -0.970031253194543992604}, {-0.219101240156869797228, -0.975702130038528544460},
    {-0.195090322016128267848, -0.980785280403230449126}, {-0.170961888760301226364, -0.985277642388941244774},
    {-0.146730474455361751659, -0.989176509964780973452}, {-0.122410675199216198499, -0.992479534598709998157},
    {-0.098017140329560601994, -0.995184726672196886245}, {-0.073564563599667423529, -0.997290456678690216136},
    {-0.049067674327418014255, -0.998795456205172392715}, {-0.024541228522912288032, -0.999698818696204220116},
    {0.0, -1.0}, {0.024541228522912288032, -0.999698818696204220116},
    {0.049067674327418014255, -0.998795456205172392715}, {0.073564563599667423529, -0.997290456678690216136},
    {0.098017140329560601994, -0.995184726672196886245}, {0.122410675199216198499, -0.992479534598709998157},
    {0.146730474455361751659, -0.989176509964780973452}, {0.170961888760301226364, -0.985277642388941244774},
    {0.195090322016128267848, -0.980785280403230449126}, {0.219101240156869797228, -0.975702130038528544460},
    {0.242980179903263889948, -0.970031253194543992604}, {0.266712757474898386325, -0.963776065795439866686},
    {0.290284677254462367636, -0.956940335732208864936}, {0.313681740398891476656, -0.949528180593036667196},
    {0.336889853392220050689, -0.941544065183020778413}, {0.359895036534988148775, -0.932992798834738887712},
    {0.382683432365089771728, -0.923879532511286756128}, {0.405241314004989870908, -0.914209755703530654635},
    {0.427555093430282094321, -0.903989293123443331586}, {0.449611329654606600046, -0.893224301195515320342},
    {0.471396736825997648556, -0.881921264348355029713}, {0.492898192229784036873, -0.870086991108711418652},
    {0.514102744193221726594, -0.857728610000272069902}, {0.534997619887097210663, -0.844853565249707073260},
    {0.555570233019602224743, -0.831469612302545237079}, {0.575808191417845300746, -0.817584813151583696505},
    {0.595699304492433343467, -0.803207531480644909807}, {0.615231590580626845485, -0.788346427626606262009},
    {0.634393284163645498215, -0.773010453362736960811}, {0.653172842953776764084, -0.757208846506484547575},
    {0.671558954847018400625, -0.740951125354959091176}, {0.689540544737066924617, -0.724247082951466920941},
    {0.707106781186547524401, -0.707106781186547524401}, {0.724247082951466920941, -0.689540544737066924617},
    {0.740951125354959091176, -0.671558954847018400625}, {0.757208846506484547575, -0.653172842953776764084},
    {0.773010453362736960811, -0.634393284163645498215}, {0.788346427626606262009, -0.615231590580626845485},
    {0.803207531480644909807, -0.595699304492433343467}, {0.817584813151583696505, -0.575808191417845300746},
    {0.831469612302545237079, -0.555570233019602224743}, {0.844853565249707073260, -0.534997619887097210663},
    {0.857728610000272069902, -0.514102744193221726594}, {0.870086991108711418652, -0.492898192229784036873},
    {0.881921264348355029713, -0.471396736825997648556}, {0.893224301195515320342, -0.449611329654606600046},
    {0.903989293123443331586, -0.427555093430282094321}, {0.914209755703530654635, -0.405241314004989870908},
    {0.923879532511286756128, -0.382683432365089771728}, {0.932992798834738887712, -0.359895036534988148775},
    {0.941544065183020778413, -0.336889853392220050689}, {0.949528180593036667196, -0.313681740398891476656},
    {0.956940335732208864936, -0.290284677254462367636}, {0.963776065795439866686, -0.266712757474898386325},
    {0.970031253194543992604, -0.242980179903263889948}, {0.975702130038528544460, -0.219101240156869797228},
    {0.980785280403230449126, -0.195090322016128267848}, {0.985277642388941244774, -0.170961888760301226364},
    {0.989176509964780973452, -0.146730474455361751659}, {0.992479534598709998157, -0.122410675199216198499},
    {0.995184726672196886245, -0.098017140329560601994}, {0.997290456678690216136, -0.073564563599667423529},
    {0.998795456205172392715, -0.049067674327418014255}, {0.999698818696204220116, -0.024541228522912288032},};

template <int L>
__device__ __forceinline__ double2 Tw(int k)
{
    return c_tw[(k & (2 * L - 1)) * (128 / L)];
}

__device__ __forceinline__ double Bf16LoF(unsigned short lo) { return (double)__uint_as_float(((unsigned int)lo) << 16); }

__device__ __forceinline__ int WaveIdx() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

__device__ __forceinline__ void WaveSync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Sub-filter range [qlo, qhi] of a pair's band [k0, k0 + len) within the
// level's raw lags [lagLo, lagHi] (and <= maxn).
__device__ __forceinline__ bool QRange(int k0, int len, int maxn, int L, int lagLo, int lagHi, int& qlo, int& qhi)
{
    const int jlo = k0 > lagLo ? k0 : lagLo;
    int jhi = k0 + len - 1;
    jhi = jhi < maxn ? jhi : maxn;
    jhi = jhi < lagHi ? jhi : lagHi;
    if (len <= 0 || jlo > jhi)
        return false;
    qlo = (jlo - 1) / L;
    qhi = (jhi - 1) / L;
    return true;
}

// In-place radix-2 complex FFT of N points held in registers (N a power of
// two, N <= 128, every index a compile-time constant once unrolled): SIGN -1
// the forward transform (e^{-2 pi i n k / N}), +1 the unnormalised inverse.
// Rounding grows as log2 N, not as N as in a direct sum.
template <int N>
constexpr int Log2() { return N <= 1 ? 0 : 1 + Log2<N / 2>(); }

template <int N, int SIGN>
__device__ __forceinline__ void FftInPlace(double (&re)[N], double (&im)[N])
{
    constexpr int LG = Log2<N>();
#pragma unroll
    for (int i = 0; i < N; ++i) {
        int j = 0;
#pragma unroll
        for (int bit = 0; bit < LG; ++bit)
            j |= ((i >> bit) & 1) << (LG - 1 - bit);
        if (j > i) {
            const double tr = re[i], ti = im[i];
            re[i] = re[j];
            im[i] = im[j];
            re[j] = tr;
            im[j] = ti;
        }
    }
#pragma unroll
    for (int st = 1; st <= LG; ++st) {
        const int len = 1 << st, half = len >> 1;
#pragma unroll
        for (int j = 0; j < half; ++j) {
            const double2 w = c_tw[j * (256 / len)];
            const double wr = w.x, wi = SIGN * w.y;
#pragma unroll
            for (int i = 0; i < N; i += len) {
                const int a = i + j, b = a + half;
                const double tr = re[b] * wr - im[b] * wi;
                const double ti = re[b] * wi + im[b] * wr;
                re[b] = re[a] - tr;
                im[b] = im[a] - ti;
                re[a] = re[a] + tr;
                im[a] = im[a] + ti;
            }
        }
    }
}

// Spectrum of a real 2L-point sequence x (x[2n] in zr[n], x[2n + 1] in
// zi[n]) through the L-point complex FFT of z = x_even + i x_odd: on return
// zr[k] + i zi[k] = X_k for k = 1 .. L - 1 and (zr[0], zi[0]) = (X_0, X_L).
template <int L>
__device__ __forceinline__ void RealFft(double (&zr)[L], double (&zi)[L])
{
    FftInPlace<L, -1>(zr, zi);
    const double z0r = zr[0], z0i = zi[0];
#pragma unroll
    for (int k = 1; k <= L / 2; ++k) {
        const int kk = L - k;
        // E = (Z_k + conj Z_{L-k}) / 2, O = (Z_k - conj Z_{L-k}) / 2i,
        // X_k = E + e^{-i pi k / L} O, X_{L-k} = conj(E) - e^{i pi k / L} conj(O)
        const double er = 0.5 * (zr[k] + zr[kk]), ei = 0.5 * (zi[k] - zi[kk]);
        const double orr = 0.5 * (zi[k] + zi[kk]), oi = -0.5 * (zr[k] - zr[kk]);
        const double2 w = c_tw[k * (128 / L)];
        const double pr = w.x * orr + w.y * oi, pi = w.x * oi - w.y * orr;  // e^{-i theta} O
        zr[k] = er + pr;
        zi[k] = ei + pi;
        if (kk != k) {
            zr[kk] = er - pr;
            zi[kk] = -(ei - pi);
        }
    }
    zr[0] = z0r + z0i;
    zi[0] = z0r - z0i;
}

// The inverse of RealFft: from (zr[0], zi[0]) = (X_0, X_L) and X_k in
// zr[k] + i zi[k], k = 1 .. L - 1, the real 2L-point sequence
// x_m = (1 / 2L) sum_k X_k e^{i pi k m / L}, x[2n] in zr[n], x[2n + 1] in zi[n].
template <int L>
__device__ __forceinline__ void RealIfft(double (&zr)[L], double (&zi)[L])
{
    const double x0 = zr[0], xl = zi[0];
#pragma unroll
    for (int k = 1; k <= L / 2; ++k) {
        const int kk = L - k;
        // E = (X_k + conj X_{L-k}) / 2, O = (X_k - conj X_{L-k}) e^{i pi k / L} / 2,
        // Z_k = E + i O; Z_{L-k} from the same pair with k and L - k swapped
        const double er = 0.5 * (zr[k] + zr[kk]), ei = 0.5 * (zi[k] - zi[kk]);
        const double dr = 0.5 * (zr[k] - zr[kk]), di = 0.5 * (zi[k] + zi[kk]);
        const double2 w = c_tw[k * (128 / L)];
        const double orr = dr * w.x - di * w.y, oi = dr * w.y + di * w.x;  // e^{i theta} D
        zr[k] = er - oi;
        zi[k] = ei + orr;
        if (kk != k) {
            // for L - k: E' = conj(E), D' = -conj(D), e^{i pi (L - k) / L} = -conj(e^{i theta})
            // so O' = conj(O) and Z_{L-k} = conj(E) + i conj(O)
            zr[kk] = er + oi;
            zi[kk] = -ei + orr;
        }
    }
    zr[0] = 0.5 * (x0 + xl);
    zi[0] = 0.5 * (x0 - xl);
    FftInPlace<L, 1>(zr, zi);
#pragma unroll
    for (int n = 0; n < L; ++n) {
        zr[n] *= 1.0 / L;
        zi[n] *= 1.0 / L;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Items of one level: one wavefront per row group; lanes over the domain's
// field elements.  With items null, count[g] = number of (field, sub-filter)
// items of group g and rowCount[g] the sum of their row counts; else they are
// written from items[itemOff[g]] in field order, sub-filters ascending, each
// with the offset of its filter spectra (L slots x its rows x 30 values) from
// hOff[g] on.  Count mode also adds the items and rows of each sub-filter q
// to qHist[q] and qHist[qcap + 1 + q] (q <= qcap).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(64) k_fft_items(PairGrid grid, const NodeRefs* __restrict__ refs,
                                                  const int* __restrict__ kb, const int* __restrict__ len, int maxn,
                                                  int L, int lagLo, int lagHi, const int* __restrict__ grpDom,
                                                  const int* __restrict__ grpRowList, const int* __restrict__ grpRows,
                                                  const long* __restrict__ itemOff, const long* __restrict__ hOff,
                                                  long* __restrict__ count, long* __restrict__ rowCount,
                                                  long* __restrict__ qHist, int qcap, FftItem* __restrict__ items)
{
    __shared__ long pre[65], prr[65];
    const int g = blockIdx.x, lane = threadIdx.x;
    const int d = grpDom[g], nr = grpRows[g];
    const int* rows = grpRowList + (long)g * kFftRows;
    const int eb = grid.eleBegin[d], n = grid.eleCount[d];
    const long domPair0 = grid.pairBase[d];
    const long base = items ? itemOff[g] : 0, hbase = items ? hOff[g] : 0;
    const long perRow = (long)L * 30;
    long total = 0, rtotal = 0;
    for (int f0 = 0; f0 < n; f0 += 64) {
        const int f = f0 + lane;
        int lo = 1 << 30, hi = -1;
        if (f < n)
            for (int r = 0; r < nr; ++r) {
                const long p = domPair0 + (long)(rows[r] - eb) * n + f;
                int a, z;
                if (QRange(kb[p], len[p], maxn, L, lagLo, lagHi, a, z)) {
                    lo = a < lo ? a : lo;
                    hi = z > hi ? z : hi;
                }
            }
        // the mask of rows whose band has sub-filter q
        auto maskOf = [&](int q) {
            unsigned long long m = 0;
            for (int r = 0; r < nr; ++r) {
                const long p = domPair0 + (long)(rows[r] - eb) * n + f;
                int a, z;
                if (QRange(kb[p], len[p], maxn, L, lagLo, lagHi, a, z) && q >= a && q <= z)
                    m |= 1ull << r;
            }
            return m;
        };
        long cnt = 0, rws = 0;  // sub-filters that some row has (the rows' ranges may leave gaps)
        for (int q = lo; q <= hi; ++q) {
            const int pc = __popcll(maskOf(q));
            cnt += pc > 0;
            rws += pc;
            if (!items && pc && q <= qcap) {
                atomicAdd(reinterpret_cast<unsigned long long*>(qHist + q), 1ull);
                atomicAdd(reinterpret_cast<unsigned long long*>(qHist + qcap + 1 + q), (unsigned long long)pc);
            }
        }
        pre[lane + 1] = cnt;
        prr[lane + 1] = rws;
        __syncthreads();
        if (lane == 0) {
            pre[0] = prr[0] = 0;
            for (int i = 1; i <= 64; ++i) {
                pre[i] += pre[i - 1];
                prr[i] += prr[i - 1];
            }
        }
        __syncthreads();
        if (items && cnt) {
            long at = base + total + pre[lane];
            long ho = hbase + (rtotal + prr[lane]) * perRow;
            const int cu = (int)refs[eb + f].hu_block, ct = (int)refs[eb + f].ht_block;
            for (int q = lo; q <= hi; ++q) {
                const unsigned long long m = maskOf(q);
                if (!m)
                    continue;
                FftItem it;
                it.f = f;
                it.q = q;
                it.cu = cu;
                it.ct = ct;
                it.mask = m;
                it.hoff = ho;
                items[at++] = it;
                ho += (long)__popcll(m) * perRow;
            }
        }
        total += pre[64];
        rtotal += prr[64];
        __syncthreads();
    }
    if (!items && lane == 0) {
        count[g] = total;
        rowCount[g] = rtotal;
    }
}

void LaunchFftItems(PairGrid grid, const NodeRefs* refs, const int* kb, const int* len, int maxn, int L, int lagLo,
                    int lagHi, int ngroups, const int* grpDom, const int* grpRowList, const int* grpRows,
                    const long* itemOff, const long* hOff, long* count, long* rowCount, long* qHist, int qcap,
                    FftItem* items, hipStream_t s)
{
    k_fft_items<<<(unsigned)ngroups, 64, 0, s>>>(grid, refs, kb, len, maxn, L, lagLo, lagHi, grpDom, grpRowList, grpRows,
                                                 itemOff, hOff, count, rowCount, qHist, qcap, items);
}

// ---------------------------------------------------------------------------
// Window spectra of one level: one thread per (column, component, load
// case).  Window beta holds F_x, x = L beta - 2L + 1 + n, n = 0 .. 2L - 1,
// formed from the raw states as (g_{x-1} + 2 g_x) + g_{x+1}, transformed by a
// real FFT in registers.  W[(((col * L + slot) * 3 + c) * 2 + ri) * B + b].
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(256) k_fft_windows(const double* __restrict__ g, long cols, int nstore, int B,
                                                     int beta, double* __restrict__ W)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= cols * 3 * B)
        return;
    const long col = idx / (3 * B);
    const int rem = (int)(idx - col * 3 * B);
    const int c = rem / B, b = rem - c * B;
    const double* gc = g + (col * nstore * 3 + c) * (long)B + b;
    auto G = [&](int m) { return (m >= 1 && m < nstore) ? gc[(long)m * 3 * B] : 0.0; };
    const int x0 = L * beta - 2 * L + 1;
    double gv[2 * L + 2];
#pragma unroll
    for (int n = 0; n < 2 * L + 2; ++n)
        gv[n] = G(x0 - 1 + n);
    double zr[L], zi[L];
#pragma unroll
    for (int n = 0; n < L; ++n) {
        zr[n] = (gv[2 * n] + 2.0 * gv[2 * n + 1]) + gv[2 * n + 2];
        zi[n] = (gv[2 * n + 1] + 2.0 * gv[2 * n + 2]) + gv[2 * n + 3];
    }
    RealFft<L>(zr, zi);
    double* out = W + (col * L * 3 + c) * 2 * (long)B + b;
#pragma unroll
    for (int k = 0; k < L; ++k) {
        out[(long)k * 6 * B] = zr[k];
        out[(long)k * 6 * B + B] = zi[k];
    }
}

void LaunchFftWindows(const double* g, long cols, int nstore, int B, int L, int beta, double* W, hipStream_t s)
{
    const long n = cols * 3 * B;
    if (n <= 0)
        return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    switch (L) {
    case 16: k_fft_windows<16><<<nb, 256, 0, s>>>(g, cols, nstore, B, beta, W); break;
    case 32: k_fft_windows<32><<<nb, 256, 0, s>>>(g, cols, nstore, B, beta, W); break;
    case 64: k_fft_windows<64><<<nb, 256, 0, s>>>(g, cols, nstore, B, beta, W); break;
    default: break;
    }
}

// ---------------------------------------------------------------------------
// Sub-filter spectra of one level: one wavefront per item (its row group
// found in itemOff); lane (r, ch) = (lane / 15, lane % 15) of the item's rows
// four at a time holds the L lags of its channel in registers (raw lags
// L q + 1 .. L q + L inside the level's range, the band and maxn; -M for the
// 9 MTS channels, G for the 6 MGS ones) and transforms them with a real FFT.
// H[hoff + (slot * rows + k) * 30 + ch * 2 + ri] for the item's k-th row.
// ---------------------------------------------------------------------------
template <typename TB, int L>
__global__ void __launch_bounds__(64) k_fft_filters(
    PairGrid grid, const long* __restrict__ off, const int* __restrict__ kb, const int* __restrict__ len,
    const TB* __restrict__ mts, const TB* __restrict__ mgs, const unsigned short* __restrict__ mtsLo,
    const unsigned short* __restrict__ mgsLo, const int* __restrict__ grpDom, const int* __restrict__ grpRowList,
    int ngroups, const long* __restrict__ itemOff, const FftItem* __restrict__ items, int maxn, int lagLo, int lagHi,
    double* __restrict__ H)
{
    // one-wavefront workgroups: an item's work varies with its rows, and a
    // wavefront that finishes frees its slot at once
    __shared__ int rowOf[64];
    const int lane = threadIdx.x & 63;
    const long it = (long)blockIdx.x;
    if (it >= itemOff[ngroups])
        return;
    int glo = 0, ghi = ngroups - 1;  // the group g with itemOff[g] <= it < itemOff[g + 1]
    while (glo < ghi) {
        const int mid = (glo + ghi + 1) >> 1;
        if (itemOff[mid] <= it)
            glo = mid;
        else
            ghi = mid - 1;
    }
    const int g = glo;
    int* rw = rowOf;
    const int d = grpDom[g];
    const int* rows = grpRowList + (long)g * kFftRows;
    const int eb = grid.eleBegin[d], n = grid.eleCount[d];
    const long domPair0 = grid.pairBase[d];
    const FftItem item = items[it];
    const int f = item.f, q = item.q;
    const unsigned long long mask = item.mask;
    const int pc = __popcll(mask);
    if ((mask >> lane) & 1)
        rw[__popcll(mask & ((1ull << lane) - 1))] = lane;
    WaveSync();
    const int kk = lane / 15, ch = lane - kk * 15;
    const int jhi = maxn < lagHi ? maxn : lagHi;
    for (int k0 = 0; k0 < pc; k0 += 4) {
        const bool row = lane < 60 && k0 + kk < pc;
        const long p = domPair0 + (long)(rows[row ? rw[k0 + kk] : rw[0]] - eb) * n + f;
        const int k0b = kb[p], lp = len[p];
        // the L taps of this lane's channel: every load issued (entry 0 where
        // the lag is outside the band), then the values kept where it is not
        const bool isM = ch < 9;
        const TB* src = isM ? mts : mgs;
        const unsigned short* srcLo = isM ? mtsLo : mgsLo;
        const int nv = isM ? 9 : 6, cv = isM ? ch : ch - 9;
        const long e0 = off[p] - k0b;
        TB raw[L];
        unsigned short lo[L];
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const int j = L * q + 1 + m;
            const bool ok = row && j >= lagLo && j <= jhi && j >= k0b && j < k0b + lp;
            const long e = ok ? e0 + j : 0;
            raw[m] = src[nv * e + cv];
            lo[m] = srcLo ? srcLo[nv * e + cv] : (unsigned short)0;
        }
        double h[L];
#pragma unroll
        for (int m = 0; m < L; ++m) {
            const int j = L * q + 1 + m;
            const bool ok = row && j >= lagLo && j <= jhi && j >= k0b && j < k0b + lp;
            double val = (double)raw[m];
            if (srcLo)
                val += Bf16LoF(lo[m]);
            h[m] = ok ? (isM ? -val : val) : 0.0;
        }
        // the 2L-point spectrum of the L taps (the second half zero)
        double zr[L], zi[L];
#pragma unroll
        for (int n2 = 0; n2 < L; ++n2) {
            zr[n2] = 2 * n2 < L ? h[2 * n2] : 0.0;
            zi[n2] = 2 * n2 + 1 < L ? h[2 * n2 + 1] : 0.0;
        }
        RealFft<L>(zr, zi);
        if (row) {
            double* out = H + item.hoff + (long)(k0 + kk) * 30 + ch * 2;
#pragma unroll
            for (int slot = 0; slot < L; ++slot) {
                double* o = out + (long)slot * pc * 30;
                o[0] = zr[slot];
                o[1] = zi[slot];
            }
        }
    }
}

template <typename TB>
void LaunchFftFilters(PairGrid grid, const long* off, const int* kb, const int* len, const TB* mts, const TB* mgs,
                      const unsigned short* mtsLo, const unsigned short* mgsLo, int ngroups, const int* grpDom,
                      const int* grpRowList, const long* itemOff, long nitems, const FftItem* items, int maxn, int L,
                      int lagLo, int lagHi, double* H, hipStream_t s)
{
    if (nitems <= 0)
        return;
    const unsigned nb = (unsigned)nitems;
#define TDBEM_FILTERS(LL)                                                                                            \
    k_fft_filters<TB, LL><<<nb, 64, 0, s>>>(grid, off, kb, len, mts, mgs, mtsLo, mgsLo, grpDom, grpRowList, ngroups, \
                                             itemOff, items, maxn, lagLo, lagHi, H)
    switch (L) {
    case 16: TDBEM_FILTERS(16); break;
    case 32: TDBEM_FILTERS(32); break;
    case 64: TDBEM_FILTERS(64); break;
    default: break;
    }
#undef TDBEM_FILTERS
}

// ---------------------------------------------------------------------------
// Spectral products: one workgroup of eight wavefronts per (row group, slot,
// 64 load cases); wavefront w owns rows 8w .. 8w + 7 of the group and skips
// the items that none of them has.  Items are loaded two ahead into a ring of
// three register sets (the window spectra of the lane's load case, column f,
// window t + 1 - q and this slot; the filter spectra of the wavefront's rows,
// at most 240 values, four per lane, read past the item's end into H's
// padding on the lanes that hold none); an item's filter spectra are written
// to the wavefront's LDS slice before its products, which read them at
// wave-uniform addresses.  The summed spectra go to
// Y[((row * L + slot) * 6 + c * 2 + ri) * B + b].  Slot 0 packs the real
// bins 0 and L, so its products are two real ones per value.
// ---------------------------------------------------------------------------
struct FftRegs {
    double xu[3][2], xt[3][2];
    double2 h0, h1;
};

template <int L, bool PACKED>
__device__ __forceinline__ void FftProducts(const FftItem* __restrict__ items, long i0, long i1, int t, int tRing,
                                            int nwr, int pace, int sh, int slot, int lane, int B, int bb,
                                            const double* __restrict__ H, const double* __restrict__ WU,
                                            const double* __restrict__ WT, long ringStrideU, long ringStrideT,
                                            double2* hs2, double (&acc)[8][3][2])
{
    const double* hs = reinterpret_cast<const double*>(hs2);
    // next item at or after `it` that has one of this wavefront's rows and
    // whose window exists in this tile (q <= t + 1)
    auto nextItem = [&](long it) {
        while (it < i1 && (items[it].q > t + 1 || ((items[it].mask >> sh) & 0xffull) == 0))
            ++it;
        return it;
    };
    auto load = [&](long it, FftRegs& x) {
        const FftItem item = items[it];
        // ring slot of window t + 1 - q: (t + 1 - q) mod nwr with 1 <= q <= nwr
        const int beta = tRing - item.q + (tRing < item.q ? nwr : 0);
        const double* wu = WU + beta * ringStrideU + ((long)item.cu * 6 * L + slot * 6) * B + bb;
        const double* wt = WT + beta * ringStrideT + ((long)item.ct * 6 * L + slot * 6) * B + bb;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri) {
                x.xu[c][ri] = wu[(long)(c * 2 + ri) * B];
                x.xt[c][ri] = wt[(long)(c * 2 + ri) * B];
            }
        const unsigned long long m = item.mask;
        const int pc = __popcll(m), below = __popcll(m & ((1ull << sh) - 1));
        const int n2 = 15 * __popcll((m >> sh) & 0xffull);  // double2 values of the wavefront's rows
        const double2* hp = reinterpret_cast<const double2*>(H + item.hoff + ((long)slot * pc + below) * 30);
        // lanes past the rows keep what they held: the products never read those values
        if (lane < n2)
            x.h0 = hp[lane];
        if (lane + 64 < n2)
            x.h1 = hp[lane + 64];
    };
    // the item's filter spectra into the wavefront's LDS slice; the wave's
    // LDS accesses run in order, so the previous item's products have read
    // theirs, and these writes land before the next products read them (the
    // wavefront-scope fences keep the compiler from moving accesses across)
    auto stage = [&](const FftRegs& x) {
        WaveSync();
        hs2[lane] = x.h0;
        hs2[lane + 64] = x.h1;
        WaveSync();
    };
    auto mac = [&](long it, const FftRegs& x) {
        const unsigned sub = (unsigned)((items[it].mask >> sh) & 0xffull);
        const int gi[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            if (!((sub >> r) & 1u))
                continue;
            const double* h = hs + __popc(sub & ((1u << r) - 1u)) * 30;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                double re = acc[r][c][0], im = acc[r][c][1];
#pragma unroll
                for (int cc = 0; cc < 3; ++cc) {
                    const double mr = h[2 * (3 * c + cc)], mi = h[2 * (3 * c + cc) + 1];
                    const double gr = h[18 + 2 * gi[c][cc]], gim = h[18 + 2 * gi[c][cc] + 1];
                    const double ur = x.xu[cc][0], ui = x.xu[cc][1];
                    const double tr = x.xt[cc][0], ti = x.xt[cc][1];
                    if (PACKED) {
                        re = fma(mr, ur, re);
                        im = fma(mi, ui, im);
                        re = fma(gr, tr, re);
                        im = fma(gim, ti, im);
                    } else {
                        re = fma(mr, ur, re);
                        re = fma(-mi, ui, re);
                        im = fma(mr, ui, im);
                        im = fma(mi, ur, im);
                        re = fma(gr, tr, re);
                        re = fma(-gim, ti, re);
                        im = fma(gr, ti, im);
                        im = fma(gim, tr, im);
                    }
                }
                acc[r][c][0] = re;
                acc[r][c][1] = im;
            }
        }
    };
    // The wavefronts of the workgroup share each item's window spectra
    // through L2: a barrier after every `pace` items of the list keeps them
    // within that many items of each other.  Every wavefront passes the same
    // nPace barriers (those of the chunks it skips on its way, and the rest at
    // the end); they order nothing in memory, so no fence goes with them.
    const int nPace = pace > 0 ? (int)((i1 - i0 + pace - 1) / pace) : 0;
    int paced = 0;
    auto paceTo = [&](long it) {
        const int chunk = pace > 0 ? (int)((it - i0) / pace) : 0;
        while (paced < chunk) {
            __builtin_amdgcn_s_barrier();
            ++paced;
        }
    };
    long cur = nextItem(i0);
    if (cur < i1) {
        FftRegs x0, x1, x2;
        long n1 = nextItem(cur + 1), n2;
        load(cur, x0);
        load(n1 < i1 ? n1 : cur, x1);
        stage(x0);
        // x0 / x1 / x2 rotate: the current item's set is staged, the next
        // one's loads are in flight and the one after is loaded before the
        // products; past the list's end the current item is loaded again
        for (;;) {
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x2);
            paceTo(cur);
            mac(cur, x0);
            if (n1 >= i1)
                break;
            stage(x1);
            cur = n1;
            n1 = n2;
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x0);
            paceTo(cur);
            mac(cur, x1);
            if (n1 >= i1)
                break;
            stage(x2);
            cur = n1;
            n1 = n2;
            n2 = n1 < i1 ? nextItem(n1 + 1) : i1;
            load(n2 < i1 ? n2 : cur, x1);
            paceTo(cur);
            mac(cur, x2);
            if (n1 >= i1)
                break;
            stage(x0);
            cur = n1;
            n1 = n2;
        }
    }
    while (paced < nPace) {
        __builtin_amdgcn_s_barrier();
        ++paced;
    }
}

template <int L>
__global__ void __launch_bounds__(512) k_conv_far_fft(
    const int* __restrict__ grpRowList, const int* __restrict__ grpRows, const long* __restrict__ itemOff,
    const FftItem* __restrict__ items, const double* __restrict__ H, const double* __restrict__ WU,
    const double* __restrict__ WT, long ringStrideU, long ringStrideT, int nwr, int t, int B, int pace,
    double* __restrict__ Y)
{
    static_assert(kFftRows == 64, "eight rows per wavefront: the filter spectra of a wavefront fill at most 240 values");
    __shared__ __align__(16) double2 Hs[8][128];
    const int slot = blockIdx.x & (L - 1), g = blockIdx.x / L;
    const int wave = WaveIdx(), lane = threadIdx.x & 63;
    const int sh = 8 * wave;
    const int b = blockIdx.y * 64 + lane;
    const bool live = b < B;
    const int bb = live ? b : B - 1;
    const int* rows = grpRowList + (long)g * kFftRows;
    const int nr = grpRows[g];
    const long i0 = itemOff[g], i1 = itemOff[g + 1];
    const int tRing = (t + 1) % nwr;
    double acc[8][3][2];
#pragma unroll
    for (int r = 0; r < 8; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            acc[r][c][0] = acc[r][c][1] = 0.0;
    if (slot == 0)
        FftProducts<L, true>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, bb, H, WU, WT, ringStrideU, ringStrideT,
                          Hs[wave], acc);
    else
        FftProducts<L, false>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, bb, H, WU, WT, ringStrideU, ringStrideT,
                           Hs[wave], acc);
    if (!live)
        return;
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        if (sh + r >= nr)
            break;
        double* y = Y + (((long)rows[sh + r] * L + slot) * 6) * B + b;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
                y[(long)(c * 2 + ri) * B] = acc[r][c][ri];
    }
}

// ---------------------------------------------------------------------------
// The same products on the FP64 matrix cores (v_mfma_f64_16x16x4f64; A (16 x
// 4): lane l holds A[l % 16][l / 16]; B (4 x 16): lane l holds
// B[l / 16][l % 16]; D (16 x 16): lane l, register q holds
// D[l / 16 + 4 q][l % 16]).  Per wavefront and item, the complex products of
// its 8 rows and 64 load cases are one real product
//   D (48 x 64) += A (48 x 12) B (12 x 64)
// with output o = 16 c + 8 ri + r (component c, real / imaginary ri, row r)
// and k = 6 field + 2 cc + wri (field U / T, component cc of the window,
// real / imaginary wri): A holds the filter values re / -im / im / re (slot
// 0: its two real bins, re . re and im . im), zero for the rows the item does
// not have; B the window spectra.  3 (c) x 3 (k steps) x 4 (16-case tiles)
// MFMAs per item; the 12 accumulator tiles stay in registers.  Each lane
// loads its own A entries (nine values of its row's filter spectra) and B
// entries one item ahead into registers, so no LDS stands between the loads
// and the MFMAs.  Items, windows and pacing as in
// k_conv_far_fft.
// ---------------------------------------------------------------------------
typedef double FftD4 __attribute__((ext_vector_type(4)));

template <int L, bool PACKED>
__device__ __forceinline__ void FftProductsMf(const FftItem* __restrict__ items, long i0, long i1, int t, int tRing,
                                              int nwr, int pace, int sh, int slot, int lane, int B, int b0,
                                              const double* __restrict__ H, const double* __restrict__ WU,
                                              const double* __restrict__ WT, long ringStrideU, long ringStrideT,
                                              FftD4 (&acc)[3][4])
{
    const int li = lane & 15, lj = lane >> 4;
    const int ri = li >> 3, r = li & 7;
    // this lane's A entries: row r, output (component c, ri), k = 4 kt + lj:
    // the value at aoff[c][kt] of the row's filter spectra, negated for bit
    // 3 c + kt of neg, zero for that bit of zero
    int aoff[3][3], neg = 0, zero = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            const int k = 4 * kt + lj, field = k / 6, cc = (k % 6) >> 1, wri = k & 1;
            const int gi = c == cc ? (c == 0 ? 0 : (c == 1 ? 3 : 5)) : (c + cc == 1 ? 1 : (c + cc == 2 ? 2 : 4));
            const int base = field == 0 ? 2 * (3 * c + cc) : 18 + 2 * gi;
            const int bit = 1 << (3 * c + kt);
            if (PACKED) {
                aoff[c][kt] = base + ri;
                zero |= ri == wri ? 0 : bit;
            } else {
                aoff[c][kt] = base + (ri ^ wri);
                neg |= (ri == 0 && wri == 1) ? bit : 0;
            }
        }
    // this lane's B entries: k = 4 kt + lj -> field, value row kk of the slot
    long boff[3];
    bool bT[3];
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
        const int k = 4 * kt + lj;
        bT[kt] = k >= 6;
        boff[kt] = ((long)slot * 6 + (k % 6)) * B;
    }
    auto nextItem = [&](long it) {
        while (it < i1 && (items[it].q > t + 1 || ((items[it].mask >> sh) & 0xffull) == 0))
            ++it;
        return it;
    };
    // window spectra (B operands) of item `it` (L2 hits: the workgroup's
    // wavefronts share them)
    auto loadW = [&](long it, double (&bw)[3][4]) {
        const FftItem item = items[it];
        const int beta = tRing - item.q + (tRing < item.q ? nwr : 0);
        const double* wu = WU + beta * ringStrideU + (long)item.cu * 6 * L * B;
        const double* wt = WT + beta * ringStrideT + (long)item.ct * 6 * L * B;
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            const double* w = (bT[kt] ? wt : wu) + boff[kt];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                int bn = b0 + nt * 16 + li;
                bn = bn < B ? bn : B - 1;
                bw[kt][nt] = w[bn];
            }
        }
    };
    // the lane's row of the item's filter spectra (A operands, read once from
    // HBM); a row the item lacks reads the wavefront's first row and its A
    // entries are zeroed
    auto loadH = [&](long it, double (&ha)[3][3]) {
        const FftItem item = items[it];
        const unsigned long long m = item.mask;
        const unsigned sub = (unsigned)((m >> sh) & 0xffull);
        const int pc = __popcll(m), below = __popcll(m & ((1ull << sh) - 1));
        const int rank = ((sub >> r) & 1u) ? __popc(sub & ((1u << r) - 1u)) : 0;
        const double* hr = H + item.hoff + ((long)slot * pc + below + rank) * 30;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int kt = 0; kt < 3; ++kt)
                ha[c][kt] = hr[aoff[c][kt]];
    };
    auto mac = [&](long it, const double (&bw)[3][4], const double (&ha)[3][3]) {
        const unsigned sub = (unsigned)((items[it].mask >> sh) & 0xffull);
        const bool have = (sub >> r) & 1u;
#pragma unroll
        for (int kt = 0; kt < 3; ++kt) {
            double a[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const int bit = 1 << (3 * c + kt);
                const double v = ha[c][kt];
                a[c] = (!have || (zero & bit)) ? 0.0 : ((neg & bit) ? -v : v);
            }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int nt = 0; nt < 4; ++nt)
                    acc[c][nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[c], bw[kt][nt], acc[c][nt], 0, 0, 0);
        }
    };
    const int nPace = pace > 0 ? (int)((i1 - i0 + pace - 1) / pace) : 0;
    int paced = 0;
    auto paceTo = [&](long it) {
        const int chunk = pace > 0 ? (int)((it - i0) / pace) : 0;
        while (paced < chunk) {
            __builtin_amdgcn_s_barrier();
            ++paced;
        }
    };
    long cur = nextItem(i0);
    if (cur < i1) {
        // two register sets: the next item's window spectra and filter values
        // are in flight while the current item's products run (36 MFMAs of
        // 64 cycles); a deeper ring does not fit the registers of two
        // wavefronts per SIMD without spilling
        double w0[3][4], w1[3][4], h0[3][3], h1[3][3];
        loadW(cur, w0);
        loadH(cur, h0);
        for (;;) {
            long n1 = nextItem(cur + 1);
            loadW(n1 < i1 ? n1 : cur, w1);
            loadH(n1 < i1 ? n1 : cur, h1);
            paceTo(cur);
            mac(cur, w0, h0);
            if (n1 >= i1)
                break;
            cur = n1;
            n1 = nextItem(cur + 1);
            loadW(n1 < i1 ? n1 : cur, w0);
            loadH(n1 < i1 ? n1 : cur, h0);
            paceTo(cur);
            mac(cur, w1, h1);
            if (n1 >= i1)
                break;
            cur = n1;
        }
    }
    while (paced < nPace) {
        __builtin_amdgcn_s_barrier();
        ++paced;
    }
}

// Launched either as workgroups of eight wavefronts, one per (row group,
// slot), or (single) as one-wavefront workgroups: block bid is wavefront
// (bid / 8) % 8 of pair (bid / 64) * 8 + bid % 8, so the eight wavefronts of
// a (row group, slot), which read the same window spectra, run on one XCD
// (block bid goes to XCD bid % 8) and are dispatched together, and a
// wavefront that runs out of items frees its slot for the next one instead
// of waiting for the rest of its workgroup; the pairs' row groups are taken
// in the given order (order[pair / L]), the groups with the most work first.
template <int L>
__global__ void __launch_bounds__(512) k_conv_far_fft_mf(
    const int* __restrict__ grpRowList, const int* __restrict__ grpRows, const long* __restrict__ itemOff,
    const FftItem* __restrict__ items, const double* __restrict__ H, const double* __restrict__ WU,
    const double* __restrict__ WT, long ringStrideU, long ringStrideT, int nwr, int t, int B, int pace,
    double* __restrict__ Y, int single, const int* __restrict__ order)
{
    const int bid = blockIdx.x;
    const int pair = single ? (bid >> 6) * 8 + (bid & 7) : bid;
    const int slot = pair & (L - 1), g = single ? order[pair / L] : pair / L;
    const int wave = single ? (bid >> 3) & 7 : WaveIdx(), lane = threadIdx.x & 63;
    const int sh = 8 * wave;
    const int b0 = blockIdx.y * 64;
    const int* rows = grpRowList + (long)g * kFftRows;
    const int nr = grpRows[g];
    const long i0 = itemOff[g], i1 = itemOff[g + 1];
    const int tRing = (t + 1) % nwr;
    FftD4 acc[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            acc[c][nt] = FftD4{0.0, 0.0, 0.0, 0.0};
    if (slot == 0)
        FftProductsMf<L, true>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, b0, H, WU, WT, ringStrideU, ringStrideT,
                            acc);
    else
        FftProductsMf<L, false>(items, i0, i1, t, tRing, nwr, pace, sh, slot, lane, B, b0, H, WU, WT, ringStrideU,
                             ringStrideT, acc);
    // D[o][n] with o = l / 16 + 4 q = 8 ri + r of component c, case 16 nt + l % 16
    const int li = lane & 15, lj = lane >> 4;
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int bn = b0 + nt * 16 + li;
            if (bn >= B)
                continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int o = lj + 4 * q, ri = o >> 3, r = o & 7;
                if (sh + r < nr)
                    Y[(((long)rows[sh + r] * L + slot) * 6 + c * 2 + ri) * B + bn] = acc[c][nt][q];
            }
        }
}

void LaunchConvFarFft(int ngroups, const int* order, const int* grpRowList, const int* grpRows, const long* itemOff,
                      const FftItem* items,
                      const double* H, const double* WU, const double* WT, long ringStrideU, long ringStrideT,
                      int nwr, int L, int t, int B, double* Y, bool mfma, int pace, hipStream_t s)
{
    dim3 g((unsigned)ngroups * (unsigned)L, (unsigned)((B + 63) / 64));
#define TDBEM_PRODUCTS(LL)                                                                                              \
    do {                                                                                                                \
        if (mfma && pace == 0)                                                                                          \
            k_conv_far_fft_mf<LL><<<dim3(8u * g.x, g.y), 64, 0, s>>>(grpRowList, grpRows, itemOff, items, H, WU, WT,    \
                                                                     ringStrideU, ringStrideT, nwr, t, B, 0, Y, 1,      \
                                                                     order);                                            \
        else if (mfma)                                                                                                  \
            k_conv_far_fft_mf<LL><<<g, 512, 0, s>>>(grpRowList, grpRows, itemOff, items, H, WU, WT, ringStrideU,        \
                                                    ringStrideT, nwr, t, B, pace, Y, 0, order);                         \
        else                                                                                                            \
            k_conv_far_fft<LL><<<g, 512, 0, s>>>(grpRowList, grpRows, itemOff, items, H, WU, WT, ringStrideU,           \
                                                 ringStrideT, nwr, t, B, pace, Y);                                      \
    } while (0)
    switch (L) {
    case 16: TDBEM_PRODUCTS(16); break;
    case 32: TDBEM_PRODUCTS(32); break;
    case 64: TDBEM_PRODUCTS(64); break;
    default: break;
    }
#undef TDBEM_PRODUCTS
}

// ---------------------------------------------------------------------------
// Inverse transforms of one level: one thread per (source row, component,
// load case); samples L .. L + Tn - 1 of the 2L-point inverse transform of
// the row's summed spectra (a real FFT in registers) into
// hfar[(i * 3ne + 3 row + c) * B + b], plus, for each of the add.n buffers
// of the same layout, its value at the same index.
// ---------------------------------------------------------------------------
template <int L>
__global__ void __launch_bounds__(256) k_fft_inverse(const double* __restrict__ Y, int ne, int B, int Tn,
                                                     double* __restrict__ hfar, FftAdds add)
{
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long)ne * 3 * B)
        return;
    const long row = idx / (3 * B);
    const int rem = (int)(idx - row * 3 * B);
    const int c = rem / B, b = rem - c * B;
    const double* y = Y + (row * L * 6 + c * 2) * (long)B + b;
    double zr[L], zi[L];
#pragma unroll
    for (int k = 0; k < L; ++k) {
        zr[k] = y[(long)k * 6 * B];
        zi[k] = y[(long)k * 6 * B + B];
    }
    RealIfft<L>(zr, zi);
#pragma unroll
    for (int i = 0; i < L; ++i) {
        if (i >= Tn)
            continue;
        const int m = L + i;
        double v = (m & 1) ? zi[m >> 1] : zr[m >> 1];
        const long at = ((long)i * 3 * ne + 3 * row + c) * B + b;
        for (int a = 0; a < add.n; ++a)
            v += add.p[a][at];
        hfar[at] = v;
    }
}

void LaunchFftInverse(const double* Y, int ne, int B, int L, int Tn, double* hfar, FftAdds add, hipStream_t s)
{
    const long n = (long)ne * 3 * B;
    if (n <= 0 || Tn <= 0)
        return;
    const unsigned nb = (unsigned)((n + 255) / 256);
    switch (L) {
    case 16: k_fft_inverse<16><<<nb, 256, 0, s>>>(Y, ne, B, Tn, hfar, add); break;
    case 32: k_fft_inverse<32><<<nb, 256, 0, s>>>(Y, ne, B, Tn, hfar, add); break;
    case 64: k_fft_inverse<64><<<nb, 256, 0, s>>>(Y, ne, B, Tn, hfar, add); break;
    default: break;
    }
}

template void LaunchFftFilters<double>(PairGrid, const long*, const int*, const int*, const double*, const double*,
                                       const unsigned short*, const unsigned short*, int, const int*, const int*,
                                       const long*, long, const FftItem*, int, int, int, int, double*, hipStream_t);
template void LaunchFftFilters<float>(PairGrid, const long*, const int*, const int*, const float*, const float*,
                                      const unsigned short*, const unsigned short*, int, const int*, const int*,
                                      const long*, long, const FftItem*, int, int, int, int, double*, hipStream_t);

}  // namespace tdbem
