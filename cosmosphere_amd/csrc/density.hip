/*
 * density.hip -- the smoothing-length Newton-Raphson density kernel, its closure, the
 * active-set compaction and the post-convergence gradient sweep.
 *
 * Restates src/density_base.cu:237-477 (density sums and NR update), :479-714 (gradient
 * sweep) and :856-886 (compaction) for gfx950.
 *
 * Mapping: one wavefront per i-leaf (bucket 64 = wave64), lane = particle of the leaf.
 * The leaf's neighbour-leaf list is consumed 64 entries at a time: each lane reads one
 * j-leaf's particle range, a wave scan of the range lengths lays the j-particles out as
 * consecutive slots, and the slots are staged 64 at a time into an LDS tile with one
 * coalesced load per lane (j-leaf particles are contiguous in Hilbert order).  Every lane
 * then sweeps the tile with broadcast LDS reads.
 *
 * Precision forms:
 *   plain  (Real = double)  the reference's FP64 arithmetic;
 *   mixed  (Real = float)   positions enter the pair loop as FP32 offsets from the i-leaf's
 *                           geometric centre (formed in FP64, rounded once per j per tile),
 *                           the kernel sums are FP32 over each tile of <= 64 neighbours and
 *                           tile sums are added with Kahan compensation; the NR update of h,
 *                           the convergence test and the closure rho(h) stay FP64.
 *
 * Semantics of the active set (src/density_base.cu:856-886): a particle iterates until it
 * converges and is never revived (FLAT_PARTICLE); in WARP_PER_LEAF mode every particle of
 * an active leaf is recomputed but only the still-active ones can leave the active set
 * (src/density_base.cu:361-364).  Active leaves are the leaves holding an active particle,
 * compacted in leaf order, which equals the reference's sort + unique of their leaves.
 */
#include "density.hpp"
#include "prims.hpp"
#include "tile.hpp"

#include <type_traits>

namespace csph
{

constexpr int kDensWaves   = 4;
constexpr int kDensThreads = kDensWaves * kWave;

// The reference's NR update, in FP64: src/density_base.cu:294-333.
__device__ inline void nrUpdate(double wsum, double gsum, double pmass, double tolh, double hfact, double& h,
                                double& rhoOut, double& gradOut, bool& conv)
{
#pragma clang fp contract(off)
    const double hi = h, hiOld = h;
    const double hi1 = 1.0 / hi, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
    const double rho_i  = wsum * kCnormk * pmass * hi31;
    const double grad_i = gsum * kCnormk * pmass * hi41;
    if (!(rho_i > 0.0))
    {
        rhoOut = 0.0; gradOut = 0.0; conv = false;
        return;
    }
    const double rhoh_i  = pmass * (hfact * hi1) * (hfact * hi1) * (hfact * hi1);
    const double funci   = rhoh_i - rho_i;
    const double dhdrhoi = -hi / (3.0 * rhoh_i);
    const double omegai  = 1.0 - dhdrhoi * grad_i;
    const double safe    = (omegai > 0.0) ? omegai : fabs(omegai + 1e-300);
    const double hraw    = hi - funci * dhdrhoi / safe;
    double hnew;
    if      (hraw > 1.2 * hi) hnew = 1.2 * hi;
    else if (hraw < 0.8 * hi) hnew = 0.8 * hi;
    else                      hnew = hraw;
    rhoOut  = rho_i;
    gradOut = grad_i;
    h       = hnew;
    conv    = fabs((hnew - hi) / hiOld) < tolh;
}

__device__ inline void blockCounters(unsigned long long* __restrict__ counters, unsigned long long v0,
                                     unsigned long long v1)
{
    __shared__ unsigned long long red[2][kDensWaves];
    v0 = waveReduceSum(v0);
    v1 = waveReduceSum(v1);
    const int w = threadIdx.x >> 6;
    if (laneId() == 0) { red[0][w] = v0; red[1][w] = v1; }
    __syncthreads();
    if (threadIdx.x == 0)
    {
        unsigned long long s0 = 0, s1 = 0;
        for (int k = 0; k < kDensWaves; ++k) { s0 += red[0][k]; s1 += red[1][k]; }
        unsigned long long* slot = counters + (blockIdx.x % kCounterSlots) * kCounterStride;
        if (s0) atomicAdd(slot + 0, s0);
        if (s1) atomicAdd(slot + 1, s1);
    }
}

template<class Real, bool PER, int MODE>
__global__ __launch_bounds__(kDensThreads) void densityKernel(DensArgs A)
{
    constexpr bool kMixed = sizeof(Real) == 4;
    // mixed: j offsets as 32-bit fixed point from the i-leaf centre; plain: FP64 positions
    using TileT = typename std::conditional<kMixed, int, double>::type;
    __shared__ int   segStart[kDensWaves][kWave];
    __shared__ int   segJ[kDensWaves][kWave];
    __shared__ TileT tx[kDensWaves][kWave], ty[kDensWaves][kWave], tz[kDensWaves][kWave];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = laneId();
    const int a = blockIdx.x * kDensWaves + w;
    unsigned long long nAcc = 0, nCand = 0;

    if (a < A.nLeaves)
    {
        const int iLeaf = A.leaves ? A.leaves[a] : a;
        const uint32_t p0 = A.layout[iLeaf], p1 = A.layout[iLeaf + 1];
        const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        DBox box{};
        if constexpr (PER) box = A.boxes[A.leafSet[iLeaf]];
        // Mixed form: offsets from the i-leaf centre c in units of range / 2^29, where range
        // = leaf half-size + R hmax bounds every i offset.  The difference of two such
        // integers is exact, so a pair separation is rounded once, to FP32, relative to
        // itself (not to the offsets).  j offsets are clamped to +-2^30: a clamped j is at
        // least `range` from every i, beyond R h_i, and differences stay inside int32.
        double c0 = 0.0, c1 = 0.0, c2 = 0.0, S = 0.0, invS2 = 0.0;
        if constexpr (kMixed)
        {
            const int node = A.leafNode[iLeaf];
            c0 = A.cx[node]; c1 = A.cy[node]; c2 = A.cz[node];
            double hm = A.hmaxLeaf[iLeaf];
            if (!isfinite(hm) || hm <= 0.0) hm = 0.0;
            const double range = fmax(fmax(A.sx[node], A.sy[node]), A.sz[node]) + kRad * hm;
            S = 536870912.0 / range;   // 2^29
            invS2 = 1.0 / (S * S);
        }
        auto toFixed = [&](double v) -> int {
            v = rint(v * S);
            v = fmin(fmax(v, -1073741824.0), 1073741824.0);   // +-2^30
            return (int)v;
        };
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const uint32_t i     = q0 + lane;
            const bool     valid = i < p1;
            const bool     doLane = valid && (MODE != 0 || A.active[i]);
            const uint64_t doMask = ballot64(doLane);
            if (doMask == 0) continue;
            const double xi = valid ? A.x[i] : 0.0, yi = valid ? A.y[i] : 0.0, zi = valid ? A.z[i] : 0.0;
            double hcur = valid ? A.h[i] : 1.0;
            const double hsi = 1.0 / (hcur * hcur);
            TileT xr, yr, zr;
            if constexpr (kMixed) { xr = toFixed(xi - c0); yr = toFixed(yi - c1); zr = toFixed(zi - c2); }
            else                  { xr = xi; yr = yi; zr = zi; }
            // q^2 = r^2 h^-2 in fixed-point units: hi part for the candidate test, the
            // remainder refines it for accepted pairs
            const double hsd = kMixed ? hsi * invS2 : hsi;
            const Real hsr = Real(hsd);
            const Real hsrLo = kMixed ? Real(hsd - (double)hsr) : Real(0);
            // Sums are FP64 in both forms: the mixed form evaluates each pair term in FP32
            // and adds it to an FP64 accumulator (full rate on this GPU, cheaper than FP32
            // compensated summation and exact to the rounding of the terms themselves).
            double sw = 0, sg = 0;
            int  acc = 0;
            uint32_t cand = 0;
            auto load = [&](int l, int j) {
                double xj = A.x[j], yj = A.y[j], zj = A.z[j];
                if constexpr (kMixed)
                {
                    xj -= c0; yj -= c1; zj -= c2;
                    if constexpr (PER)
                    {
                        if (fabs(xj) > 0.5 * box.l[0]) xj -= box.l[0] * copysign(1.0, xj);
                        if (fabs(yj) > 0.5 * box.l[1]) yj -= box.l[1] * copysign(1.0, yj);
                        if (fabs(zj) > 0.5 * box.l[2]) zj -= box.l[2] * copysign(1.0, zj);
                    }
                    tx[w][l] = toFixed(xj); ty[w][l] = toFixed(yj); tz[w][l] = toFixed(zj);
                }
                else
                {
                    tx[w][l] = xj; ty[w][l] = yj; tz[w][l] = zj;
                }
            };
            auto consume = [&](int nt) {
#pragma unroll 4
                for (int k = 0; k < nt; ++k)
                {
                    Real dx, dy, dz;
                    if constexpr (kMixed)
                    {
                        dx = (Real)(xr - tx[w][k]); dy = (Real)(yr - ty[w][k]); dz = (Real)(zr - tz[w][k]);
                    }
                    else
                    {
                        dx = xr - tx[w][k]; dy = yr - ty[w][k]; dz = zr - tz[w][k];
                        if constexpr (PER)
                        {
                            if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
                            if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
                            if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
                        }
                    }
                    const Real r2 = dx * dx + dy * dy + dz * dz;
                    Real q2 = r2 * hsr;
                    if (q2 < Real(kRad2))
                    {
                        Real q;
                        if constexpr (kMixed)
                        {
                            q2 = fmaf(r2, hsrLo, q2);
                            q = __builtin_amdgcn_sqrtf(q2);
                        }
                        else
                        {
                            q = sqrt(q2);
                        }
                        Real wk, dwk;
                        kernWdW<Real>(q2, q, wk, dwk);
                        sw += (double)wk;                               // list order, as the reference
                        sg += (double)(-q * dwk - Real(3) * wk);
                        ++acc;
                    }
                }
                cand += (uint32_t)nt;
            };
            forEachTile(segStart[w], segJ[w], A.layout, A.jList, lo, hi, load, consume);

            const double wsum = sw;
            const double gsum = sg;
            if (doLane)
            {
                nAcc += (unsigned long long)acc;
                nCand += cand;
                if constexpr (MODE == 2)
                {
#pragma clang fp contract(off)
                    const double hi1 = 1.0 / hcur, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
                    A.rho[i]   = wsum * kCnormk * A.pmass * hi31;
                    A.gradh[i] = gsum * kCnormk * A.pmass * hi41;
                }
                else
                {
                    double r, g;
                    bool   conv;
                    nrUpdate(wsum, gsum, A.pmass, A.tolh, A.hfact, hcur, r, g, conv);
                    A.rho[i]   = r;
                    A.gradh[i] = g;
                    A.h[i]     = hcur;
                    if (MODE == 0) A.active[i] = conv ? 0 : 1;
                    else if (conv) A.active[i] = 0;
                }
            }
        }
    }
    blockCounters(A.counters, nAcc, nCand);
}

// Active-leaf compaction: flag[a] = leaf a still holds an active particle; counts the
// active particles.
__global__ __launch_bounds__(256) void leafFlagKernel(const int* __restrict__ leaves, int nLeaves,
                                                      const uint32_t* __restrict__ layout,
                                                      const uint8_t* __restrict__ active, uint32_t* __restrict__ flag,
                                                      unsigned long long* __restrict__ nActive)
{
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (a >= nLeaves) return;
    const int l = leaves[a];
    const uint32_t p0 = layout[l], p1 = layout[l + 1];
    unsigned long long c = 0;
    for (uint32_t p = p0 + laneId(); p < p1; p += kWave) c += active[p] ? 1 : 0;
    c = waveReduceSum(c);
    if (laneId() == 0)
    {
        flag[a] = c ? 1u : 0u;
        if (c) atomicAdd(nActive + (blockIdx.x % kCounterSlots) * kCounterStride, c);
    }
}

__global__ void leafScatterKernel(const int* __restrict__ leaves, int nLeaves, const uint32_t* __restrict__ flag,
                                  const uint32_t* __restrict__ pos, int* __restrict__ out)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < nLeaves && flag[a]) out[pos[a]] = leaves[a];
}

__global__ void initActiveKernel(uint8_t* __restrict__ active, int nAlive, int n, double* __restrict__ rho,
                                 double* __restrict__ gradh)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    active[i] = i < nAlive ? 1 : 0;
    rho[i] = 0.0;
    gradh[i] = 0.0;
}

__global__ void iotaInt(int* __restrict__ v, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

__global__ void maxHKernel(const double* __restrict__ h, int n, unsigned long long* __restrict__ out)
{
    double m = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = fmax(m, h[i]);
    m = waveReduceMax(m);
    if (laneId() == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));   // m >= 0
}

// ---------------------------------------------------------------------------------------
// Gradient sweep: src/density_base.cu:529-714, one wavefront per leaf, all leaves.
// ---------------------------------------------------------------------------------------
__device__ inline double xiLimiter(const double g[9])
{
#pragma clang fp contract(off)
    const double dvxdx = (float)g[0], dvxdy = (float)g[1], dvxdz = (float)g[2];
    const double dvydx = (float)g[3], dvydy = (float)g[4], dvydz = (float)g[5];
    const double dvzdx = (float)g[6], dvzdy = (float)g[7], dvzdz = (float)g[8];
    const double divv = dvxdx + dvydy + dvzdz;
    const double cx = dvzdy - dvydz, cy = dvxdz - dvzdx, cz = dvydx - dvxdy;
    const double mdiv = fmax(-divv, 0.0);
    const double fac = mdiv * mdiv;
    const double tr = cx * cx + cy * cy + cz * cz;
    return (fac + tr > 2.220446049250313e-16) ? fac / (fac + tr) : 1.0;
}

template<bool PER>
__global__ __launch_bounds__(kDensThreads) void gradientKernel(GradArgs A)
{
    __shared__ int    segStart[kDensWaves][kWave];
    __shared__ int    segJ[kDensWaves][kWave];
    __shared__ double tj[kDensWaves][9][kWave];
    const int w = threadIdx.x >> 6, lane = laneId();
    const int a = blockIdx.x * kDensWaves + w;
    if (a >= A.nLeaves) return;
    const int iLeaf = a;
    const uint32_t p0 = A.layout[iLeaf], p1 = A.layout[iLeaf + 1];
    const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
    DBox box{};
    if constexpr (PER) box = A.boxes[A.leafSet[iLeaf]];
    const double pmass = A.pmass;
    for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
    {
        const uint32_t i = q0 + lane;
        const bool valid = i < p1;
        const double xi = valid ? A.x[i] : 0.0, yi = valid ? A.y[i] : 0.0, zi = valid ? A.z[i] : 0.0;
        const double vxi = valid ? A.vx[i] : 0.0, vyi = valid ? A.vy[i] : 0.0, vzi = valid ? A.vz[i] : 0.0;
        const double axi = valid ? A.ax[i] : 0.0, ayi = valid ? A.ay[i] : 0.0, azi = valid ? A.az[i] : 0.0;
        const double hi_ = valid ? A.h[i] : 1.0;
        const double hsi = 1.0 / (hi_ * hi_);
        double rhoi = 0, gradhi = 0, divs = 0, dv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, da[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        double rxx = 0, rxy = 0, rxz = 0, ryy = 0, ryz = 0, rzz = 0;
        auto load = [&](int l, int j) {
            tj[w][0][l] = A.x[j];  tj[w][1][l] = A.y[j];  tj[w][2][l] = A.z[j];
            tj[w][3][l] = A.vx[j]; tj[w][4][l] = A.vy[j]; tj[w][5][l] = A.vz[j];
            tj[w][6][l] = A.ax[j]; tj[w][7][l] = A.ay[j]; tj[w][8][l] = A.az[j];
        };
        auto consume = [&](int nt) {
            for (int k = 0; k < nt; ++k)
            {
                double dx = xi - tj[w][0][k], dy = yi - tj[w][1][k], dz = zi - tj[w][2][k];
                if constexpr (PER)
                {
                    if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
                    if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
                    if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
                }
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double q2 = r2 * hsi;
                if (q2 >= kRad2) continue;
                const double q = sqrt(q2);
                double wk, dwk;
                kernWdW<double>(q2, q, wk, dwk);
                rhoi += wk;
                gradhi += -q * dwk - 3.0 * wk;
                const double rij = sqrt(r2), rij1 = 1.0 / (rij + 2.220446049250313e-16);
                const double rg = rij1 * dwk;
                const double rx = dx * rg * pmass, ry = dy * rg * pmass, rz = dz * rg * pmass;
                const double dvx = vxi - tj[w][3][k], dvy = vyi - tj[w][4][k], dvz = vzi - tj[w][5][k];
                divs += dvx * rx + dvy * ry + dvz * rz;
                dv[0] += dvx * rx; dv[1] += dvx * ry; dv[2] += dvx * rz;
                dv[3] += dvy * rx; dv[4] += dvy * ry; dv[5] += dvy * rz;
                dv[6] += dvz * rx; dv[7] += dvz * ry; dv[8] += dvz * rz;
                const double dax = axi - tj[w][6][k], day = ayi - tj[w][7][k], daz = azi - tj[w][8][k];
                da[0] += dax * rx; da[1] += dax * ry; da[2] += dax * rz;
                da[3] += day * rx; da[4] += day * ry; da[5] += day * rz;
                da[6] += daz * rx; da[7] += daz * ry; da[8] += daz * rz;
                rxx -= dx * rx; rxy -= dx * ry; rxz -= dx * rz;
                ryy -= dy * ry; ryz -= dy * rz; rzz -= dz * rz;
            }
        };
        forEachTile(segStart[w], segJ[w], A.layout, A.jList, lo, hi, load, consume);
        if (!valid) continue;

        const double hi1 = 1.0 / hi_, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
        const double rho_i = rhoi * kCnormk * pmass * hi31;
        const double grad_i = gradhi * kCnormk * pmass * hi41;
        A.rho[i] = rho_i;
        A.gradh[i] = grad_i;
        if (!(rho_i > 0.0))
        {
            A.divv[i] = 0.0; A.ddivvdt[i] = 0.0; A.xi[i] = 1.0;
            continue;
        }
        const double omega = 1.0 + (hi_ / (3.0 * rho_i)) * grad_i;
        const double oinv = (omega > 0.0) ? (1.0 / omega) : 1.0;
        const double term = kCnormk * oinv * hi41 / rho_i;
        const double denom = rxx * ryy * rzz + 2.0 * rxy * rxz * ryz - rxx * ryz * ryz - ryy * rxz * rxz - rzz * rxy * rxy;
        const double rm0 = ryy * rzz - ryz * ryz, rm1 = rxz * ryz - rzz * rxy, rm2 = rxy * ryz - rxz * ryy;
        const double rm3 = rzz * rxx - rxz * rxz, rm4 = rxy * rxz - rxx * ryz, rm5 = rxx * ryy - rxy * rxy;
        A.divv[i] = -divs * term;
        double g[9], div_a;
        if (fabs(denom) > 2.2250738585072014e-308)
        {
            const double dd = 1.0 / denom;
#pragma unroll
            for (int row = 0; row < 3; ++row)
            {
                const double sx = dv[3 * row], sy = dv[3 * row + 1], sz = dv[3 * row + 2];
                g[3 * row]     = -(sx * rm0 + sy * rm1 + sz * rm2) * dd;
                g[3 * row + 1] = -(sx * rm1 + sy * rm3 + sz * rm4) * dd;
                g[3 * row + 2] = -(sx * rm2 + sy * rm4 + sz * rm5) * dd;
            }
            double gd = 0.0;
#pragma unroll
            for (int row = 0; row < 3; ++row)
            {
                const double sx = da[3 * row], sy = da[3 * row + 1], sz = da[3 * row + 2];
                const double gax = (sx * rm0 + sy * rm1 + sz * rm2) * dd;
                const double gay = (sx * rm1 + sy * rm3 + sz * rm4) * dd;
                const double gaz = (sx * rm2 + sy * rm4 + sz * rm5) * dd;
                gd += (row == 0) ? gax : ((row == 1) ? gay : gaz);
            }
            div_a = -gd;
        }
        else
        {
#pragma unroll
            for (int c = 0; c < 9; ++c) g[c] = -dv[c] * term;
            div_a = -term * (da[0] + da[4] + da[8]);
        }
        A.xi[i] = xiLimiter(g);
        A.ddivvdt[i] = div_a - (g[0] * g[0] + g[4] * g[4] + g[8] * g[8] + 2.0 * (g[1] * g[3] + g[2] * g[6] + g[5] * g[7]));
    }
}

__global__ void deadGradientKernel(int first, int n, double* __restrict__ divv, double* __restrict__ xi,
                                   double* __restrict__ ddivvdt)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    divv[i] = 0.0; ddivvdt[i] = 0.0; xi[i] = 1.0;
}

// ---------------------------------------------------------------------------------------
// Host wrappers
// ---------------------------------------------------------------------------------------
void launchDensity(const DensArgs& a, int precision, int mode, bool periodic, hipStream_t st)
{
    if (a.nLeaves <= 0) return;
    const int grid = iceil(a.nLeaves, kDensWaves);
#define CSPH_DENS(R, P, M) densityKernel<R, P, M><<<grid, kDensThreads, 0, st>>>(a)
    if (precision == kMixed)
    {
        if (periodic) { if (mode == 0) CSPH_DENS(float, true, 0); else if (mode == 1) CSPH_DENS(float, true, 1); else CSPH_DENS(float, true, 2); }
        else          { if (mode == 0) CSPH_DENS(float, false, 0); else if (mode == 1) CSPH_DENS(float, false, 1); else CSPH_DENS(float, false, 2); }
    }
    else
    {
        if (periodic) { if (mode == 0) CSPH_DENS(double, true, 0); else if (mode == 1) CSPH_DENS(double, true, 1); else CSPH_DENS(double, true, 2); }
        else          { if (mode == 0) CSPH_DENS(double, false, 0); else if (mode == 1) CSPH_DENS(double, false, 1); else CSPH_DENS(double, false, 2); }
    }
#undef CSPH_DENS
    LAUNCH_CHECK();
}

void launchGradients(const GradArgs& a, bool periodic, hipStream_t st)
{
    if (a.nLeaves > 0)
    {
        const int grid = iceil(a.nLeaves, kDensWaves);
        if (periodic) gradientKernel<true><<<grid, kDensThreads, 0, st>>>(a);
        else          gradientKernel<false><<<grid, kDensThreads, 0, st>>>(a);
        LAUNCH_CHECK();
    }
}

void launchDeadGradients(int first, int n, double* divv, double* xi, double* ddivvdt, hipStream_t st)
{
    if (n > first)
    {
        deadGradientKernel<<<iceil(n - first, 256), 256, 0, st>>>(first, n, divv, xi, ddivvdt);
        LAUNCH_CHECK();
    }
}

int counterSlots() { return kCounterSlots; }
int counterStride() { return kCounterStride; }

void launchLeafFlags(const int* leaves, int nLeaves, const uint32_t* layout, const uint8_t* active, uint32_t* flag,
                     unsigned long long* nActive, hipStream_t st)
{
    if (nLeaves <= 0) return;
    leafFlagKernel<<<iceil(nLeaves, 4), 256, 0, st>>>(leaves, nLeaves, layout, active, flag, nActive);
    LAUNCH_CHECK();
}

void launchLeafScatter(const int* leaves, int nLeaves, const uint32_t* flag, const uint32_t* pos, int* out,
                       hipStream_t st)
{
    if (nLeaves <= 0) return;
    leafScatterKernel<<<iceil(nLeaves, 256), 256, 0, st>>>(leaves, nLeaves, flag, pos, out);
    LAUNCH_CHECK();
}

void launchInitActive(uint8_t* active, int nAlive, int n, double* rho, double* gradh, hipStream_t st)
{
    if (n <= 0) return;
    initActiveKernel<<<iceil(n, 256), 256, 0, st>>>(active, nAlive, n, rho, gradh);
    LAUNCH_CHECK();
}

void launchIota(int* v, int n, hipStream_t st)
{
    if (n <= 0) return;
    iotaInt<<<iceil(n, 256), 256, 0, st>>>(v, n);
    LAUNCH_CHECK();
}

void launchMaxH(const double* h, int n, unsigned long long* out, hipStream_t st)
{
    if (n <= 0) return;
    maxHKernel<<<std::min(iceil(n, 256), 1024), 256, 0, st>>>(h, n, out);
    LAUNCH_CHECK();
}

} // namespace csph
