/*
 * density.hip -- the gradient sweep after convergence and small helpers of the solve (the
 * Newton-Raphson pass itself is in nr.hip).
 *
 * Restates src/density_base.cu:479-714 (gradient sweep: rho and d(rho)/d(h) at the final h,
 * div v, the velocity-gradient matrix through the inverse of the r-tensor, the xi limiter and
 * d(div v)/dt) for gfx950: in FP64 as the reference, or in the mixed form (FP32 pair terms,
 * FP64 rho sums; gradientKernel<..., true>).
 *
 * Mapping: one wavefront per leaf, its particles up to 64 at a time, g = floor(64 / c) lanes
 * per particle (below); the particles of a leaf with a long list one wavefront each
 * (SweepItems).  The leaf's neighbour-leaf list is consumed 64 entries at a time: each
 * lane reads one j-leaf's particle range, a wave scan of the range lengths lays the
 * j-particles out as consecutive slots, and the slots are staged into an LDS tile with one
 * coalesced load per lane (j-leaf particles are contiguous in Hilbert order).  Every lane then
 * sweeps its slots of the tile with broadcast LDS reads.
 */
#include "density.hpp"
#include "prims.hpp"
#include "tile.hpp"

#include <type_traits>

namespace csph
{

constexpr int kDensWaves   = 4;
constexpr int kDensThreads = kDensWaves * kWave;

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

// List candidates of a pass, one wavefront per active leaf position (grid-stride): the particle
// counts of the list's leaves summed over the lanes, times the leaf's evaluated particles.
__global__ __launch_bounds__(256) void listCandKernel(const PassCtl* __restrict__ ctl, int nFix,
                                                      const int* __restrict__ leaves,
                                                      const uint32_t* __restrict__ jOffset,
                                                      const uint32_t* __restrict__ jCount,
                                                      const int* __restrict__ jList,
                                                      const uint32_t* __restrict__ layout,
                                                      const uint8_t* __restrict__ state,
                                                      unsigned long long* __restrict__ out)
{
    if (ctl && ctl->skip) return;
    const int n = ctl ? (int)ctl->nAL : nFix;
    const int lane = laneId();
    const int wave0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, waves = (gridDim.x * blockDim.x) >> 6;
    unsigned long long s1 = 0, s2 = 0;
    for (int a = wave0; a < n; a += waves)
    {
        const uint32_t lo = jOffset[a], hi = jCount ? lo + jCount[a] : jOffset[a + 1];
        unsigned long long sj = 0;
        for (uint32_t e = lo + lane; e < hi; e += kWave)
        {
            const int jl = jList[e];
            sj += layout[jl + 1] - layout[jl];
        }
        sj = waveReduceSum(sj);
        const int leaf = leaves ? leaves[a] : a;
        unsigned long long n1 = 0, n2 = 0;
        for (uint32_t p = layout[leaf] + lane; p < layout[leaf + 1]; p += kWave)
        {
            const uint8_t st = state ? state[p] : 1;
            n1 += st == 1;
            n2 += st == 2;
        }
        s1 += sj * waveReduceSum(n1);
        s2 += sj * waveReduceSum(n2);
    }
    if (lane == 0 && s1) atomicAdd(out, s1);
    if (lane == 0 && s2) atomicAdd(out + 1, s2);
}

void launchListCand(const PassCtl* ctl, int n, const int* leaves, const uint32_t* jOffset, const uint32_t* jCount,
                    const int* jList, const uint32_t* layout, const uint8_t* state, unsigned long long* out,
                    hipStream_t st)
{
    // a fixed grid when the count is on the device; grid-stride beyond it
    const int blocks = ctl ? 2048 : std::max(1, std::min(2048, iceil(n, 4)));
    listCandKernel<<<blocks, 256, 0, st>>>(ctl, n, leaves, jOffset, jCount, jList, layout, state, out);
    LAUNCH_CHECK();
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

// One wavefront per leaf, the leaf's particles up to 64 at a time.  With c particles in a
// chunk, each takes g = floor(64 / c) lanes that split every tile (lane s of a group sweeps
// slots s, s + g, ...), then a group reduction closes the particle: small leaves keep the
// whole wavefront busy.  A list entry is staged only if its leaf's particles (their tight box)
// lie within the support R h_i of some particle of the chunk: the reference's list reaches R hmax
// around the whole leaf box, most of which no particle's own support touches, and a leaf the
// test drops holds no pair with q < R.  Of a kept leaf, a particle is staged only if it lies
// within the largest support of the chunk's bounding box; the survivors fill the tiles in list
// order (forEachTileCompact), so the pair sums are unchanged (for g = 1 even their order is:
// the list order).
//
// MIXED: pair separations, velocity and acceleration differences are FP64 differences rounded
// once to FP32; r^2, q, the root, the kernel brackets (kernVD) and the 25 gradient sums are FP32;
// the W and q D brackets of rho and d(rho)/d(h) are summed in FP64, with the FP32 rounding of
// 1/h^2 taken out per particle as in the NR pass.  The 3x3 solve and the limiter are FP64.
//
// The sweep waits on its gathers more than on arithmetic: the open-box mixed instance runs five
// wavefronts per SIMD (96 VGPRs, 24 spilled), 2% faster on the 16M-particle mixture than four
// without spills; the others keep four.
template<bool PER, bool MIXED>
__global__ __launch_bounds__(kDensThreads) __attribute__((amdgpu_waves_per_eu(MIXED && !PER ? 5 : 4))) void gradientKernel(GradArgs A)
{
    using G = std::conditional_t<MIXED, float, double>;   // the gradient sums
    __shared__ int    segStart[kDensWaves][kWave];
    __shared__ int    segJ[kDensWaves][kWave];
    __shared__ int    tileHead[kDensWaves][kWave];   // first tile position of each list entry (tileOwner)
    __shared__ double tj[kDensWaves][9][kWave];
    __shared__ double cp[kDensWaves][4][kWave];   // the chunk's particles: x, y, z, (R h)^2 slack
    __shared__ int    jq[kDensWaves][2 * kWave];  // candidates that passed the particle test
    __shared__ double cb[kDensWaves][8];          // the chunk's box: centre, half-sizes, (R hmax)^2 slack
    const int w = threadIdx.x >> 6, lane = laneId();
    int a;
    uint32_t p0, p1;
    if (!A.items.resolve(blockIdx.x * kDensWaves + w, A.nLeaves, A.layout, a, p0, p1)) return;
    const int iLeaf = a;
    const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
    const DBox box = A.boxes[A.leafSet[iLeaf]];
    const double pmass = A.pmass;
    for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
    {
        const int c = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kWave, p1 - q0));
        const int g = __builtin_amdgcn_readfirstlane(kWave / c);
        const int grp = lane / g, sub = lane - grp * g;
        const bool has = grp < c;
        const uint32_t i = q0 + (uint32_t)(has ? grp : 0);
        const double xi = A.x[i], yi = A.y[i], zi = A.z[i];
        const double vxi = A.vx[i], vyi = A.vy[i], vzi = A.vz[i];
        const double axi = A.ax[i], ayi = A.ay[i], azi = A.az[i];
        const double hi_ = A.h[i];
        const double hsi = 1.0 / (hi_ * hi_), hinvi = 1.0 / hi_;
        const float hsr = float(hsi), hinvf = float(hinvi);
        const double eps = (hsi - (double)hsr) / (double)hsr;   // hsr (1 + eps) = 1/h^2
        {
            const uint32_t ic = q0 + (uint32_t)(lane < c ? lane : 0);
            const double xc = A.x[ic], yc = A.y[ic], zc = A.z[ic], hc = A.h[ic];
            if (lane < c)
            {
                cp[w][0][lane] = xc; cp[w][1][lane] = yc; cp[w][2][lane] = zc;
                cp[w][3][lane] = (kRad * hc) * (kRad * hc) * kCullSlack;
            }
            // the chunk's bounding box and largest reach, kept in LDS (registers are scarce here)
            double c3[3], s3[3], rm;
            waveBox(lane < c, xc, yc, zc, kRad * hc, c3, s3, rm);
            if (lane == 0)
            {
                cb[w][0] = c3[0]; cb[w][1] = c3[1]; cb[w][2] = c3[2];
                cb[w][3] = s3[0]; cb[w][4] = s3[1]; cb[w][5] = s3[2];
                cb[w][6] = rm * rm * kCullSlack;
            }
        }
        waveLdsSync();
        double rhoi = 0, gradhi = 0;   // plain: sum W, sum -q dW/dq - 3W; mixed: sum V, sum q D
        G divs = 0, dv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, da[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        G rxx = 0, rxy = 0, rxz = 0, ryy = 0, ryz = 0, rzz = 0;
        auto keep = [&](int jl) {
            for (int k = 0; k < c; ++k)
                if (pointLeafGap2<PER>(cp[w][0][k], cp[w][1][k], cp[w][2][k], A.lbox, A.lboxN, jl, box) <= cp[w][3][k])
                    return true;
            return false;
        };
        // of a kept leaf, a particle only if it lies within the largest support of the chunk's box
        auto keepJ = [&](int j) {
            const double c3[3] = {cb[w][0], cb[w][1], cb[w][2]}, s3[3] = {cb[w][3], cb[w][4], cb[w][5]};
            return pointBoxGap2<PER>(A.x[j], A.y[j], A.z[j], c3, s3, box) <= cb[w][6];
        };
        // all nine loads in flight before the first LDS write (one memory round trip per tile;
        // left to itself the scheduler pairs each two loads with their write)
        auto load = [&](int l, int j) {
            const double f[9] = {A.x[j], A.y[j], A.z[j], A.vx[j], A.vy[j], A.vz[j], A.ax[j], A.ay[j], A.az[j]};
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < 9; ++q) tj[w][q][l] = f[q];
        };
        auto sep = [&](int k, double& dx, double& dy, double& dz) {
            dx = xi - tj[w][0][k]; dy = yi - tj[w][1][k]; dz = zi - tj[w][2][k];
            if constexpr (PER)
            {
                if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
                if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
                if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
            }
        };
        auto pairMixed = [&](int k) {
            double dxd, dyd, dzd;
            sep(k, dxd, dyd, dzd);
            const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
            const float r2 = dx * dx + dy * dy + dz * dz;
            const float q2 = r2 * hsr;
            if (!(q2 < float(kRad2))) return;
            const float q = __builtin_amdgcn_sqrtf(q2);
            float V, qD;
            kernVD<float>(q, q2, V, qD);
            rhoi += (double)V;
            gradhi += (double)qD;
            // dW/dq / r = -kGp D / r = -kGp q D / (r^2 / h); the self pair (r = 0) has dx = 0
            const float rg = r2 > 0.0f ? float(-kGp) * qD * __builtin_amdgcn_rcpf(r2 * hinvf) : 0.0f;
            const float pm = (float)pmass;
            const float rx = dx * rg * pm, ry = dy * rg * pm, rz = dz * rg * pm;
            const float dvx = (float)(vxi - tj[w][3][k]), dvy = (float)(vyi - tj[w][4][k]);
            const float dvz = (float)(vzi - tj[w][5][k]);
            divs += dvx * rx + dvy * ry + dvz * rz;
            dv[0] += dvx * rx; dv[1] += dvx * ry; dv[2] += dvx * rz;
            dv[3] += dvy * rx; dv[4] += dvy * ry; dv[5] += dvy * rz;
            dv[6] += dvz * rx; dv[7] += dvz * ry; dv[8] += dvz * rz;
            const float dax = (float)(axi - tj[w][6][k]), day = (float)(ayi - tj[w][7][k]);
            const float daz = (float)(azi - tj[w][8][k]);
            da[0] += dax * rx; da[1] += dax * ry; da[2] += dax * rz;
            da[3] += day * rx; da[4] += day * ry; da[5] += day * rz;
            da[6] += daz * rx; da[7] += daz * ry; da[8] += daz * rz;
            rxx -= dx * rx; rxy -= dx * ry; rxz -= dx * rz;
            ryy -= dy * ry; ryz -= dy * rz; rzz -= dz * rz;
        };
        auto pair = [&](int k) {
            if constexpr (MIXED)
            {
                pairMixed(k);
                return;
            }
            double dx, dy, dz;
            sep(k, dx, dy, dz);
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double q2 = r2 * hsi;
            if (q2 >= kRad2) return;
            // one root: q = r / h (within an ulp of sqrt(q^2)); 1 / (r + eps) by the hardware
            // reciprocal and two Newton steps (within an ulp of the division)
            const double rij = sqrt(r2), q = rij * hinvi;
            double wk, dwk;
            kernWdW<double>(q2, q, wk, dwk);
            rhoi += wk;
            gradhi += -q * dwk - 3.0 * wk;
            const double den = rij + 2.220446049250313e-16;
            double rij1 = __builtin_amdgcn_rcp(den);   // as recip() in force.hip
            rij1 = fma(rij1, fma(-den, rij1, 1.0), rij1);
            rij1 = fma(rij1, fma(-den, rij1, 1.0), rij1);
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
        };
        // lane sub of a group takes slots sub, sub + g, ... below nt (g = 1: slots 0 .. nt - 1,
        // the list order, as the reference), in two phases: the support test of every slot into
        // a bit mask, then the pair terms of the accepted slots in slot order.  A wavefront runs
        // the pair terms as often as its busiest lane accepts a pair, not once per candidate
        // (at the final h about 1 candidate in 8 is a pair); each lane's sums are unchanged.
        auto consume = [&](int nt) {
            uint64_t accept = 0;
            int t = 0;
            for (int k = sub; k < nt; k += g, ++t)
            {
                double dx, dy, dz;
                sep(k, dx, dy, dz);
                bool in;
                if constexpr (MIXED)
                {
                    const float fx = (float)dx, fy = (float)dy, fz = (float)dz;
                    in = (fx * fx + fy * fy + fz * fz) * hsr < float(kRad2);
                }
                else
                {
                    in = (dx * dx + dy * dy + dz * dz) * hsi < kRad2;
                }
                if (in) accept |= 1ull << t;
            }
            while (accept)
            {
                const int tt = __builtin_ctzll(accept);
                accept &= accept - 1;
                pair(sub + tt * g);
            }
        };
        forEachTileCompact(segStart[w], segJ[w], tileHead[w], jq[w], A.layout, A.jList, lo, hi, keep, keepJ, load, consume);
        // group sums: a tree over lanes sub .. g - 1 of each group (lane sub = 0 ends with them)
        for (int off = 1; off < g; off <<= 1)
        {
            auto add = [&](auto& v) {
                const auto o = __shfl_down(v, off, kWave);
                if (sub + off < g) v += o;
            };
            add(rhoi); add(gradhi); add(divs);
#pragma unroll
            for (int f = 0; f < 9; ++f) { add(dv[f]); add(da[f]); }
            add(rxx); add(rxy); add(rxz); add(ryy); add(ryz); add(rzz);
        }
        if (!has || sub != 0) continue;

        const double hi1 = 1.0 / hi_, hi31 = hi1 * hi1 * hi1, hi41 = hi31 * hi1;
        double wsum = rhoi, gsum = gradhi;
        if constexpr (MIXED)
        {
            // brackets to sums, the q of every pair short by (1 + eps/2): as particleSums in nr.hip
            const double v = rhoi - (0.5 * eps) * (kGp / kWw) * gradhi;
            wsum = kWw * v;
            gsum = kGp * gradhi - kGw * v;
        }
        const double rho_i = wsum * kCnormk * pmass * hi31;
        const double grad_i = gsum * kCnormk * pmass * hi41;
        A.rho[i] = rho_i;
        A.gradh[i] = grad_i;
        if (!(rho_i > 0.0))
        {
            A.divv[i] = 0.0; A.ddivvdt[i] = 0.0; A.xi[i] = 1.0;
            continue;
        }
        const double omega = 1.0 + (hi_ / (3.0 * rho_i)) * grad_i;
        const double oinv = (omega > 0.0) ? (1.0 / omega) : 1.0;
        // the sums in FP64 from here on (the mixed form accumulates them in FP32)
        const double Rxx = rxx, Rxy = rxy, Rxz = rxz, Ryy = ryy, Ryz = ryz, Rzz = rzz, Divs = divs;
        double Dv[9], Da[9];
#pragma unroll
        for (int f = 0; f < 9; ++f) { Dv[f] = dv[f]; Da[f] = da[f]; }
        const double term = kCnormk * oinv * hi41 / rho_i;
        const double denom = Rxx * Ryy * Rzz + 2.0 * Rxy * Rxz * Ryz - Rxx * Ryz * Ryz - Ryy * Rxz * Rxz - Rzz * Rxy * Rxy;
        const double rm0 = Ryy * Rzz - Ryz * Ryz, rm1 = Rxz * Ryz - Rzz * Rxy, rm2 = Rxy * Ryz - Rxz * Ryy;
        const double rm3 = Rzz * Rxx - Rxz * Rxz, rm4 = Rxy * Rxz - Rxx * Ryz, rm5 = Rxx * Ryy - Rxy * Rxy;
        A.divv[i] = -Divs * term;
        double gm[9], div_a;
        if (fabs(denom) > 2.2250738585072014e-308)
        {
            const double dd = 1.0 / denom;
#pragma unroll
            for (int row = 0; row < 3; ++row)
            {
                const double sx = Dv[3 * row], sy = Dv[3 * row + 1], sz = Dv[3 * row + 2];
                gm[3 * row]     = -(sx * rm0 + sy * rm1 + sz * rm2) * dd;
                gm[3 * row + 1] = -(sx * rm1 + sy * rm3 + sz * rm4) * dd;
                gm[3 * row + 2] = -(sx * rm2 + sy * rm4 + sz * rm5) * dd;
            }
            double gd = 0.0;
#pragma unroll
            for (int row = 0; row < 3; ++row)
            {
                const double sx = Da[3 * row], sy = Da[3 * row + 1], sz = Da[3 * row + 2];
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
            for (int f = 0; f < 9; ++f) gm[f] = -Dv[f] * term;
            div_a = -term * (Da[0] + Da[4] + Da[8]);
        }
        A.xi[i] = xiLimiter(gm);
        A.ddivvdt[i] =
            div_a - (gm[0] * gm[0] + gm[4] * gm[4] + gm[8] * gm[8] + 2.0 * (gm[1] * gm[3] + gm[2] * gm[6] + gm[5] * gm[7]));
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
void launchGradients(const GradArgs& a, bool periodic, bool mixed, hipStream_t st)
{
    if (a.nLeaves > 0)
    {
        const int grid = iceil(a.items.count(a.nLeaves), kDensWaves);
        if (mixed)
        {
            if (periodic) gradientKernel<true, true><<<grid, kDensThreads, 0, st>>>(a);
            else          gradientKernel<false, true><<<grid, kDensThreads, 0, st>>>(a);
        }
        else
        {
            if (periodic) gradientKernel<true, false><<<grid, kDensThreads, 0, st>>>(a);
            else          gradientKernel<false, false><<<grid, kDensThreads, 0, st>>>(a);
        }
        LAUNCH_CHECK();
    }
}

// Sweep items: leaf a with more than `entries` list entries and more than one particle gets
// one item per particle.
__global__ void sweepCountKernel(int nLeaves, const uint32_t* __restrict__ jOffset, const uint32_t* __restrict__ layout,
                                 int entries, uint8_t* __restrict__ flag, uint32_t* __restrict__ count)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nLeaves) return;
    const uint32_t np = layout[a + 1] - layout[a];
    const bool heavy = entries > 0 && jOffset[a + 1] - jOffset[a] > (uint32_t)entries && np > 1u;
    flag[a] = heavy ? 1 : 0;
    count[a] = heavy ? np : 0u;
}

__global__ void sweepFillKernel(int nLeaves, const uint32_t* __restrict__ layout, const uint8_t* __restrict__ flag,
                                const uint32_t* __restrict__ off, int2* __restrict__ items)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= nLeaves || !flag[a]) return;
    const uint32_t p0 = layout[a], p1 = layout[a + 1];
    for (uint32_t p = p0; p < p1; ++p) items[off[a] + (p - p0)] = make_int2(a, (int)p);
}

SweepItems buildSweepItems(State& s, int nLeaves, int entries, hipStream_t st)
{
    SweepItems it;
    s.sweepHeavyItems = 0;
    if (nLeaves <= 0 || entries <= 0) return it;
    uint8_t* flag = s.sweepFlag.get<uint8_t>((size_t)nLeaves);
    uint32_t* count = s.sweepCount.get<uint32_t>((size_t)nLeaves + 1);
    uint32_t* off = s.sweepOff.get<uint32_t>((size_t)nLeaves + 1);
    sweepCountKernel<<<iceil(nLeaves, 256), 256, 0, st>>>(nLeaves, s.jOffset.as<uint32_t>(), s.layout.as<uint32_t>(),
                                                          entries, flag, count);
    LAUNCH_CHECK();
    scanExclusive<uint32_t>(count, off, nLeaves, s.scanTmp, nullptr, st);
    uint32_t total = 0;
    HIP_CHECK(hipMemcpyAsync(&total, off + nLeaves, sizeof(total), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (total == 0) return it;
    int2* items = s.sweepItems.get<int2>(total);
    sweepFillKernel<<<iceil(nLeaves, 256), 256, 0, st>>>(nLeaves, s.layout.as<uint32_t>(), flag, off, items);
    LAUNCH_CHECK();
    it.nHeavy = (int)total;
    it.heavy = items;
    it.flag = flag;
    s.sweepHeavyItems = (int)total;
    return it;
}

void launchDeadGradients(int first, int n, double* divv, double* xi, double* ddivvdt, hipStream_t st)
{
    if (n > first)
    {
        deadGradientKernel<<<iceil(n - first, 256), 256, 0, st>>>(first, n, divv, xi, ddivvdt);
        LAUNCH_CHECK();
    }
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
