/*
 * force.hip -- per-particle force factors and the SPH pressure + artificial viscosity pair
 * sums over the symmetric neighbour-leaf lists.
 *
 * Restates src/force.cu:92-137 (forcePrepKernel) and :145-383 (sphForceKernelJList) for
 * gfx950: one wavefront per i-leaf over its symmetric list.  Only x, y, z and h of the
 * j-particles' fields go through an LDS tile: each is staged once per list entry with
 * coalesced loads and read by every particle of the chunk that evaluates it.
 */
#include "density.hpp"
#include "tile.hpp"

namespace csph
{

constexpr int kForceWaves   = 4;
constexpr int kForceThreads = kForceWaves * kWave;

__global__ __launch_bounds__(256) void forcePrepKernel(const double* __restrict__ h, const double* __restrict__ gradh,
                                                       const double* __restrict__ pro2, double* __restrict__ hsqinv,
                                                       double* __restrict__ hinv, double* __restrict__ rhoh,
                                                       double* __restrict__ rho1, double* __restrict__ grkfac,
                                                       double* __restrict__ pres, double* __restrict__ auterm,
                                                       double* __restrict__ divfac, int n, double pmass, double alphau,
                                                       double hfact)
{
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i];
    if (!(hi > 0.0))
    {
        hsqinv[i] = 0.0; hinv[i] = 0.0; rhoh[i] = 0.0; rho1[i] = 0.0; grkfac[i] = 0.0;
        pres[i] = 0.0; auterm[i] = 0.0; divfac[i] = 0.0;
        return;
    }
    const double hs = 1.0 / (hi * hi), h4 = hs * hs, hfoh = hfact / hi;
    const double rho = pmass * hfoh * hfoh * hfoh;
    const double dhdrho = -hi / (3.0 * rho);
    const double omega = 1.0 - dhdrho * gradh[i];
    const double oinv = (omega > 0.0) ? (1 / omega) : 1.0;
    hsqinv[i] = hs;
    hinv[i]   = 1.0 / hi;
    rhoh[i]   = rho;
    rho1[i]   = 1.0 / rho;
    grkfac[i] = h4 * kCnormk * oinv;
    pres[i]   = pro2[i] * rho * rho;
    auterm[i] = 0.5 * pmass * (1.0 / rho) * alphau;
    divfac[i] = kCnormk * oinv * h4 / rho;
}

// 1/x for x > 0 by the hardware reciprocal and two Newton steps: within an ulp of the division,
// at a third of its instructions.
__device__ inline double recip(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
}

// One wavefront per i-leaf, the leaf's particles up to 64 at a time, g = floor(64 / c) lanes
// per particle splitting every tile (lane s of a group takes slots s, s + g, ...), then a group
// reduction; small leaves keep the wavefront busy.  A list entry is staged only if its leaf's
// tight box lies within R max(h_i, hmax of the j-leaf) of some particle of the chunk, and of its
// leaf a particle j only if it lies within R max(h_j, max h of the chunk) of the chunk's bounding
// box: a dropped leaf or particle holds no pair inside either support, so the pair sums are
// unchanged (for g = 1 even their order; the survivors fill the tiles in list order).  On the
// 16M-particle Plummer mixture the particle test halves the candidates (11.1 to 5.5 per pair).
template<bool PER, bool DISC, bool FULL>
__global__ __launch_bounds__(kForceThreads) void forceKernel(ForceArgs A)
{
    __shared__ int    segStart[kForceWaves][kWave];
    __shared__ int    segJ[kForceWaves][kWave];
    __shared__ int    tileHead[kForceWaves][kWave];   // first tile position of each list entry (tileOwner)
    // the tile: every j-field a pair reads, staged once per tile slot with coalesced loads
    // (a j is accepted by several particles of the chunk: reading its fields per pair would
    // be one scattered load per field per pair)
    enum { X, Y, Z, H, VX, VY, VZ, HS, HI, GK, RH, R1, P2, CS, AL, U, PR, AU, NF };
    __shared__ double tf[kForceWaves][NF][kWave];
    __shared__ double cp[kForceWaves][4][kWave];   // the chunk's particles: x, y, z, R h
    __shared__ int    jq[kForceWaves][2 * kWave];  // candidates that passed the particle test
    const int w = threadIdx.x >> 6, lane = laneId();
    int a;
    uint32_t p0, p1;
    unsigned long long nAcc = 0, nCand = 0;
    if (A.items.resolve(blockIdx.x * kForceWaves + w, A.nLeaves, A.layout, a, p0, p1))
    {
        const int iLeaf = a;
        const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        const DBox box = A.boxes[A.leafSet[iLeaf]];
        const double pmass = A.pmass, beta = A.beta;
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const int c = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kWave, p1 - q0));
            const int g = __builtin_amdgcn_readfirstlane(kWave / c);
            const int grp = lane / g, sub = lane - grp * g;
            const bool has = grp < c;
            const uint32_t i = q0 + (uint32_t)(has ? grp : 0);
            const double xi = A.x[i], yi = A.y[i], zi = A.z[i], hi_ = A.h[i];
            const double vxi = A.vx[i], vyi = A.vy[i], vzi = A.vz[i];
            const double hsi = A.hsqinv[i], rhoi = A.rhoh[i], gki = A.grkfac[i], hinvi = A.hinv[i];
            const double pro2i = A.pro2[i], vwi = A.cs[i], ali = A.alpha[i], eni = A.u[i];
            const double rho1i = A.rho1[i], pri = A.pres[i], aui = A.auterm[i];
            double cbc[3], cbs[3], rChunk;   // the chunk's bounding box and largest reach R h
            {
                const uint32_t ic = q0 + (uint32_t)(lane < c ? lane : 0);
                const double xc = A.x[ic], yc = A.y[ic], zc = A.z[ic], rc = kRad * A.h[ic];
                if (lane < c)
                {
                    cp[w][0][lane] = xc; cp[w][1][lane] = yc; cp[w][2][lane] = zc;
                    cp[w][3][lane] = rc;
                }
                waveBox(lane < c, xc, yc, zc, rc, cbc, cbs, rChunk);
            }
            waveLdsSync();
            double ix = 0, iy = 0, iz = 0, jx = 0, jy = 0, jz = 0, f4s = 0, vmax = 0, divs = 0;
            int acc = 0;
            uint32_t cand = 0;
            auto keep = [&](int jl) {
                double hj = A.hmaxLeaf[jl];
                const double rj = (isfinite(hj) && hj > 0.0) ? kRad * hj : 0.0;
                for (int k = 0; k < c; ++k)
                {
                    const double r = fmax(cp[w][3][k], rj);
                    if (pointLeafGap2<PER>(cp[w][0][k], cp[w][1][k], cp[w][2][k], A.lbox, A.lboxN, jl, box) <=
                        r * r * kCullSlack)
                        return true;
                }
                return false;
            };
            auto keepJ = [&](int j) {
                const double r = fmax(rChunk, kRad * A.h[j]);
                return pointBoxGap2<PER>(A.x[j], A.y[j], A.z[j], cbc, cbs, box) <= r * r * kCullSlack;
            };
            auto load = [&](int l, int j) {
                tf[w][X][l] = A.x[j]; tf[w][Y][l] = A.y[j]; tf[w][Z][l] = A.z[j]; tf[w][H][l] = A.h[j];
                tf[w][VX][l] = A.vx[j]; tf[w][VY][l] = A.vy[j]; tf[w][VZ][l] = A.vz[j];
                tf[w][HS][l] = A.hsqinv[j]; tf[w][HI][l] = A.hinv[j]; tf[w][GK][l] = A.grkfac[j];
                tf[w][RH][l] = A.rhoh[j]; tf[w][R1][l] = A.rho1[j]; tf[w][P2][l] = A.pro2[j];
                tf[w][CS][l] = A.cs[j]; tf[w][AL][l] = A.alpha[j]; tf[w][U][l] = A.u[j];
                tf[w][PR][l] = A.pres[j]; tf[w][AU][l] = A.auterm[j];
            };
            auto sep = [&](int k, double& dx, double& dy, double& dz) {
                dx = xi - tf[w][X][k]; dy = yi - tf[w][Y][k]; dz = zi - tf[w][Z][k];
                if constexpr (PER)
                {
                    if (fabs(dx) > 0.5 * box.l[0]) dx = dx - box.l[0] * copysign(1.0, dx);
                    if (fabs(dy) > 0.5 * box.l[1]) dy = dy - box.l[1] * copysign(1.0, dy);
                    if (fabs(dz) > 0.5 * box.l[2]) dz = dz - box.l[2] * copysign(1.0, dz);
                }
            };
            // the pair test: 0 < r < R max(h_i, h_j)
            auto accepts = [&](int k) {
                double dx, dy, dz;
                sep(k, dx, dy, dz);
                const double dr2 = dx * dx + dy * dy + dz * dz;
                const double hj = tf[w][H][k];
                const double rmax2 = kRad2 * fmax(hi_ * hi_, hj * hj) * (1.0 + 1e-10);
                return (dr2 > 0.0) && (dr2 < rmax2);
            };
            auto pair = [&](int k) {
                double dx, dy, dz;
                sep(k, dx, dy, dz);
                const double dr2 = dx * dx + dy * dy + dz * dz;
                const double hj = tf[w][H][k];
                const double rmax2 = kRad2 * fmax(hi_ * hi_, hj * hj) * (1.0 + 1e-10);
                if (!((dr2 > 0.0) && (dr2 < rmax2))) return;
                const double qi2 = dr2 * hsi, qj2 = dr2 * tf[w][HS][k];
                const double dr = sqrt(dr2), rij1 = recip(dr);
                const double rx = dx * rij1, ry = dy * rij1, rz = dz * rij1;
                const double dvx = vxi - tf[w][VX][k], dvy = vyi - tf[w][VY][k], dvz = vzi - tf[w][VZ][k];
                const double projv = dvx * rx + dvy * ry + dvz * rz;
                const double rhoj = tf[w][RH][k], rho1j = tf[w][R1][k], pro2j = tf[w][P2][k];
                const double vwj = tf[w][CS][k], alj = tf[w][AL][k], enj = tf[w][U][k], prj = tf[w][PR][k];
                const double auj = tf[w][AU][k];
                const double vsigi = fmax(vwi - beta * projv, 0.0);
                const double vsigavi = fmax(ali * vwi - beta * projv, 0.0);
                const double vsigj = fmax(vwj - beta * projv, 0.0);
                const double vsigavj = fmax(alj * vwj - beta * projv, 0.0);
                const double pvmax = fmax(vsigi, vsigj);
                double qr2i = 0.0, qr2j = 0.0;
                const double denij = eni - enj;
                const double vsigu = sqrt(fabs(pri - prj) * (2.0 * recip(rhoi + rhoj)));
                if constexpr (DISC)
                {
                    const double hori = hi_ * rij1, horj = hj * rij1;
                    if (projv < 0.0)
                    {
                        qr2i = -0.5 * rho1i * (ali * vwi - beta * projv) * hori * projv;
                        qr2j = -0.5 * rho1j * (alj * vwj - beta * projv) * horj * projv;
                    }
                    else
                    {
                        qr2i = -0.5 * rho1i * ali * vwi * hori * projv;
                        qr2j = -0.5 * rho1j * alj * vwj * horj * projv;
                    }
                }
                else if (projv < 0.0)
                {
                    qr2i = -0.5 * rho1i * vsigavi * projv;
                    qr2j = -0.5 * rho1j * vsigavj * projv;
                }
                bool counted = false;
                if (qi2 < kRad2)
                {
                    counted = true;
                    vmax = fmax(vmax, pvmax);
                    double wk, dwk;
                    const double q = dr * hinvi;
                    kernWdW<double>(qi2, q, wk, dwk);
                    divs += pmass * dwk * projv;
                    const double gk = dwk * gki;
                    const double gp = pmass * (pro2i + qr2i) * gk;
                    ix += -gp * rx; iy += -gp * ry; iz += -gp * rz;
                    const double pdvt = pmass * pro2i * projv * gk;
                    const double diss = DISC ? -0.5 * pmass * rho1i * ali * vwi * (hi_ * rij1) * projv * projv * gk
                                             : pmass * qr2i * projv * gk;
                    const double dend = vsigu * denij * aui * gk;
                    if constexpr (FULL) { f4s += pdvt; f4s += diss; }
                    else
                    {
                        if (A.pdv) f4s += pdvt;
                        if (A.shock) f4s += diss;
                    }
                    f4s += dend;
                }
                if (qj2 < kRad2)
                {
                    counted = true;
                    vmax = fmax(vmax, pvmax);
                    double wk, dwk;
                    const double q = dr * tf[w][HI][k];
                    kernWdW<double>(qj2, q, wk, dwk);
                    const double gk = dwk * tf[w][GK][k];
                    const double gp = pmass * (pro2j + qr2j) * gk;
                    jx -= gp * rx; jy -= gp * ry; jz -= gp * rz;
                    f4s += vsigu * denij * auj * gk;
                }
                acc += counted ? 1 : 0;
            };
            // lane sub of a group takes slots sub, sub + g, ... below nt (g = 1: the list order):
            // the pair test of every slot into a bit mask, then the pair terms of the accepted
            // slots in slot order, so that a wavefront runs the pair terms as often as its
            // busiest lane accepts a pair rather than once per candidate
            auto consume = [&](int nt) {
                if (sub == 0) cand += (uint32_t)nt;
                uint64_t accept = 0;
                int t = 0;
                for (int k = sub; k < nt; k += g, ++t)
                    if (accepts(k)) accept |= 1ull << t;
                while (accept)
                {
                    const int tt = __builtin_ctzll(accept);
                    accept &= accept - 1;
                    pair(sub + tt * g);
                }
            };
            forEachTileCompact(segStart[w], segJ[w], tileHead[w], jq[w], A.layout, A.jList, lo, hi, keep, keepJ, load, consume);
            // group sums (and the max of vsigmax): a tree over lanes sub .. g - 1 of each group
            for (int off = 1; off < g; off <<= 1)
            {
                auto add = [&](double& v) {
                    const double o = __shfl_down(v, off, kWave);
                    if (sub + off < g) v += o;
                };
                add(ix); add(iy); add(iz); add(jx); add(jy); add(jz); add(f4s); add(divs);
                const double om = __shfl_down(vmax, off, kWave);
                if (sub + off < g) vmax = fmax(vmax, om);
                const int oa = __shfl_down(acc, off, kWave);
                if (sub + off < g) acc += oa;
            }
            if (has && sub == 0)
            {
                A.fx[i] = ix + jx; A.fy[i] = iy + jy; A.fz[i] = iz + jz;
                A.f4[i] = f4s; A.vsig[i] = vmax;
                A.divv[i] = -divs * A.divfac[i];
                nAcc += (unsigned long long)acc;
                nCand += cand;
            }
        }
    }
    nAcc = waveReduceSum(nAcc);
    nCand = waveReduceSum(nCand);
    if (laneId() == 0 && (nAcc | nCand))
    {
        unsigned long long* slot = A.counters + ((blockIdx.x * kForceWaves + w) % kCounterSlots) * kCounterStride;
        atomicAdd(slot + 0, nAcc);
        atomicAdd(slot + 1, nCand);
    }
}

// Mixed form of forceKernel: the same lists, culling, lanes and pair terms, with the pair
// algebra in FP32.  Positions, velocities and u enter as FP64 differences rounded once to FP32
// (a difference of two FP32-rounded values could lose all its digits); r^2, the support tests,
// the square roots and reciprocals, the signal speeds, the viscosity and both kernel gradients
// are FP32, from FP32 copies of the per-particle factors; each pair's terms go into FP32
// partials that are added into FP64 accumulators once per tile (at most 64 pairs per lane).
// Lanes keep their own particle's slots (force_rounds=0); forceKernelMixedRounds, the default,
// spreads each tile's accepted pairs evenly over the lanes.
template<bool PER, bool DISC, bool FULL>
__global__ __launch_bounds__(kForceThreads) __attribute__((amdgpu_waves_per_eu(4))) void forceKernelMixed(ForceArgs A)
{
    __shared__ int    segStart[kForceWaves][kWave];
    __shared__ int    segJ[kForceWaves][kWave];
    __shared__ int    tileHead[kForceWaves][kWave];   // first tile position of each list entry (tileOwner)
    // the tile: FP64 what enters a difference, FP32 the factors, all staged once per slot
    enum { X, Y, Z, VX, VY, VZ, U, ND };
    enum { H, HS, HI, GK, RH, R1, P2, CS, AL, PR, AU, NS };
    __shared__ double td[kForceWaves][ND][kWave];
    __shared__ float  ts[kForceWaves][NS][kWave];
    __shared__ double cp[kForceWaves][4][kWave];   // the chunk's particles: x, y, z, R h
    __shared__ int    jq[kForceWaves][2 * kWave];  // candidates that passed the particle test
    const int w = threadIdx.x >> 6, lane = laneId();
    const size_t S = A.f32Stride;
    auto F = [&](int k, uint32_t idx) { return A.f32[(size_t)k * S + idx]; };
    int a;
    uint32_t p0, p1;
    unsigned long long nAcc = 0, nCand = 0;
    if (A.items.resolve(blockIdx.x * kForceWaves + w, A.nLeaves, A.layout, a, p0, p1))
    {
        const int iLeaf = a;
        const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        const DBox box = A.boxes[A.leafSet[iLeaf]];
        const float pm = (float)A.pmass, beta = (float)A.beta;
        const float R2 = (float)kRad2;
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const int c = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kWave, p1 - q0));
            const int g = __builtin_amdgcn_readfirstlane(kWave / c);
            const int grp = lane / g, sub = lane - grp * g;
            const bool has = grp < c;
            const uint32_t i = q0 + (uint32_t)(has ? grp : 0);
            const double xi = A.x[i], yi = A.y[i], zi = A.z[i];
            const double vxi = A.vx[i], vyi = A.vy[i], vzi = A.vz[i], eni = A.u[i];
            const float hif = (float)A.h[i];
            const float hsi = F(0, i), hinvi = F(1, i), gki = F(2, i), rhoi = F(3, i), rho1i = F(4, i);
            const float pro2i = F(5, i), vwi = F(6, i), ali = F(7, i), pri = F(8, i), aui = F(9, i);
            double cbc[3], cbs[3], rChunk;   // the chunk's bounding box and largest reach R h
            {
                const uint32_t ic = q0 + (uint32_t)(lane < c ? lane : 0);
                const double xc = A.x[ic], yc = A.y[ic], zc = A.z[ic], rc = kRad * A.h[ic];
                if (lane < c)
                {
                    cp[w][0][lane] = xc; cp[w][1][lane] = yc; cp[w][2][lane] = zc;
                    cp[w][3][lane] = rc;
                }
                waveBox(lane < c, xc, yc, zc, rc, cbc, cbs, rChunk);
            }
            waveLdsSync();
            double ix = 0, iy = 0, iz = 0, jx = 0, jy = 0, jz = 0, f4s = 0, divs = 0;
            float pix = 0, piy = 0, piz = 0, pjx = 0, pjy = 0, pjz = 0, pf4 = 0, pdiv = 0, vmax = 0;
            int acc = 0;
            uint32_t cand = 0;
            auto keep = [&](int jl) {
                double hj = A.hmaxLeaf[jl];
                const double rj = (isfinite(hj) && hj > 0.0) ? kRad * hj : 0.0;
                for (int k = 0; k < c; ++k)
                {
                    const double r = fmax(cp[w][3][k], rj);
                    if (pointLeafGap2<PER>(cp[w][0][k], cp[w][1][k], cp[w][2][k], A.lbox, A.lboxN, jl, box) <=
                        r * r * kCullSlack)
                        return true;
                }
                return false;
            };
            auto keepJ = [&](int j) {
                const double r = fmax(rChunk, kRad * A.h[j]);
                return pointBoxGap2<PER>(A.x[j], A.y[j], A.z[j], cbc, cbs, box) <= r * r * kCullSlack;
            };
            auto load = [&](int l, int j) {
                td[w][X][l] = A.x[j]; td[w][Y][l] = A.y[j]; td[w][Z][l] = A.z[j];
                td[w][VX][l] = A.vx[j]; td[w][VY][l] = A.vy[j]; td[w][VZ][l] = A.vz[j]; td[w][U][l] = A.u[j];
                ts[w][H][l] = (float)A.h[j];
                ts[w][HS][l] = F(0, j); ts[w][HI][l] = F(1, j); ts[w][GK][l] = F(2, j); ts[w][RH][l] = F(3, j);
                ts[w][R1][l] = F(4, j); ts[w][P2][l] = F(5, j); ts[w][CS][l] = F(6, j); ts[w][AL][l] = F(7, j);
                ts[w][PR][l] = F(8, j); ts[w][AU][l] = F(9, j);
            };
            auto sep = [&](int k, float& dx, float& dy, float& dz) {
                double dxd = xi - td[w][X][k], dyd = yi - td[w][Y][k], dzd = zi - td[w][Z][k];
                if constexpr (PER)
                {
                    if (fabs(dxd) > 0.5 * box.l[0]) dxd = dxd - box.l[0] * copysign(1.0, dxd);
                    if (fabs(dyd) > 0.5 * box.l[1]) dyd = dyd - box.l[1] * copysign(1.0, dyd);
                    if (fabs(dzd) > 0.5 * box.l[2]) dzd = dzd - box.l[2] * copysign(1.0, dzd);
                }
                dx = (float)dxd; dy = (float)dyd; dz = (float)dzd;
            };
            // the pair test: 0 < r < R max(h_i, h_j)
            auto accepts = [&](int k) {
                float dx, dy, dz;
                sep(k, dx, dy, dz);
                const float dr2 = dx * dx + dy * dy + dz * dz;
                const float hj = ts[w][H][k];
                const float rmax2 = R2 * fmaxf(hif * hif, hj * hj) * (1.0f + 1e-6f);
                return (dr2 > 0.0f) && (dr2 < rmax2);
            };
            auto pair = [&](int k) {
                float dx, dy, dz;
                sep(k, dx, dy, dz);
                const float dr2 = dx * dx + dy * dy + dz * dz;
                const float hj = ts[w][H][k];
                const float rmax2 = R2 * fmaxf(hif * hif, hj * hj) * (1.0f + 1e-6f);
                if (!((dr2 > 0.0f) && (dr2 < rmax2))) return;
                const float qi2 = dr2 * hsi, qj2 = dr2 * ts[w][HS][k];
                const float dr = __builtin_amdgcn_sqrtf(dr2), rij1 = __builtin_amdgcn_rcpf(dr);
                const float rx = dx * rij1, ry = dy * rij1, rz = dz * rij1;
                const float dvx = (float)(vxi - td[w][VX][k]), dvy = (float)(vyi - td[w][VY][k]);
                const float dvz = (float)(vzi - td[w][VZ][k]);
                const float projv = dvx * rx + dvy * ry + dvz * rz;
                const float rhoj = ts[w][RH][k], rho1j = ts[w][R1][k], pro2j = ts[w][P2][k];
                const float vwj = ts[w][CS][k], alj = ts[w][AL][k], prj = ts[w][PR][k], auj = ts[w][AU][k];
                const float vsigi = fmaxf(vwi - beta * projv, 0.0f);
                const float vsigavi = fmaxf(ali * vwi - beta * projv, 0.0f);
                const float vsigj = fmaxf(vwj - beta * projv, 0.0f);
                const float vsigavj = fmaxf(alj * vwj - beta * projv, 0.0f);
                const float pvmax = fmaxf(vsigi, vsigj);
                float qr2i = 0.0f, qr2j = 0.0f;
                const float denij = (float)(eni - td[w][U][k]);
                const float vsigu = __builtin_amdgcn_sqrtf(fabsf(pri - prj) * (2.0f * __builtin_amdgcn_rcpf(rhoi + rhoj)));
                if constexpr (DISC)
                {
                    const float hori = hif * rij1, horj = hj * rij1;
                    if (projv < 0.0f)
                    {
                        qr2i = -0.5f * rho1i * (ali * vwi - beta * projv) * hori * projv;
                        qr2j = -0.5f * rho1j * (alj * vwj - beta * projv) * horj * projv;
                    }
                    else
                    {
                        qr2i = -0.5f * rho1i * ali * vwi * hori * projv;
                        qr2j = -0.5f * rho1j * alj * vwj * horj * projv;
                    }
                }
                else if (projv < 0.0f)
                {
                    qr2i = -0.5f * rho1i * vsigavi * projv;
                    qr2j = -0.5f * rho1j * vsigavj * projv;
                }
                bool counted = false;
                if (qi2 < R2)
                {
                    counted = true;
                    vmax = fmaxf(vmax, pvmax);
                    float wk, dwk;
                    const float q = dr * hinvi;
                    kernWdW<float>(qi2, q, wk, dwk);
                    pdiv += pm * dwk * projv;
                    const float gk = dwk * gki;
                    const float gp = pm * (pro2i + qr2i) * gk;
                    pix -= gp * rx; piy -= gp * ry; piz -= gp * rz;
                    const float pdvt = pm * pro2i * projv * gk;
                    const float diss = DISC ? -0.5f * pm * rho1i * ali * vwi * (hif * rij1) * projv * projv * gk
                                            : pm * qr2i * projv * gk;
                    const float dend = vsigu * denij * aui * gk;
                    if constexpr (FULL) { pf4 += pdvt; pf4 += diss; }
                    else
                    {
                        if (A.pdv) pf4 += pdvt;
                        if (A.shock) pf4 += diss;
                    }
                    pf4 += dend;
                }
                if (qj2 < R2)
                {
                    counted = true;
                    vmax = fmaxf(vmax, pvmax);
                    float wk, dwk;
                    const float q = dr * ts[w][HI][k];
                    kernWdW<float>(qj2, q, wk, dwk);
                    const float gk = dwk * ts[w][GK][k];
                    const float gp = pm * (pro2j + qr2j) * gk;
                    pjx -= gp * rx; pjy -= gp * ry; pjz -= gp * rz;
                    pf4 += vsigu * denij * auj * gk;
                }
                acc += counted ? 1 : 0;
            };
            auto flush = [&]() {
                ix += (double)pix; iy += (double)piy; iz += (double)piz;
                jx += (double)pjx; jy += (double)pjy; jz += (double)pjz;
                f4s += (double)pf4; divs += (double)pdiv;
                pix = piy = piz = pjx = pjy = pjz = pf4 = pdiv = 0.0f;
            };
            auto consume = [&](int nt) {   // in two phases, as forceKernel's
                if (sub == 0) cand += (uint32_t)nt;
                uint64_t accept = 0;
                int t = 0;
                for (int k = sub; k < nt; k += g, ++t)
                    if (accepts(k)) accept |= 1ull << t;
                while (accept)
                {
                    const int tt = __builtin_ctzll(accept);
                    accept &= accept - 1;
                    pair(sub + tt * g);
                }
                flush();
            };
            forEachTileCompact(segStart[w], segJ[w], tileHead[w], jq[w], A.layout, A.jList, lo, hi, keep, keepJ, load, consume);
            double vm = (double)vmax;
            for (int off = 1; off < g; off <<= 1)
            {
                auto add = [&](double& v) {
                    const double o = __shfl_down(v, off, kWave);
                    if (sub + off < g) v += o;
                };
                add(ix); add(iy); add(iz); add(jx); add(jy); add(jz); add(f4s); add(divs);
                const double om = __shfl_down(vm, off, kWave);
                if (sub + off < g) vm = fmax(vm, om);
                const int oa = __shfl_down(acc, off, kWave);
                if (sub + off < g) acc += oa;
            }
            if (has && sub == 0)
            {
                A.fx[i] = ix + jx; A.fy[i] = iy + jy; A.fz[i] = iz + jz;
                A.f4[i] = f4s; A.vsig[i] = vm;
                A.divv[i] = -divs * A.divfac[i];
                nAcc += (unsigned long long)acc;
                nCand += cand;
            }
        }
    }
    nAcc = waveReduceSum(nAcc);
    nCand = waveReduceSum(nCand);
    if (laneId() == 0 && (nAcc | nCand))
    {
        unsigned long long* slot = A.counters + ((blockIdx.x * kForceWaves + w) % kCounterSlots) * kCounterStride;
        atomicAdd(slot + 0, nAcc);
        atomicAdd(slot + 1, nCand);
    }
}

// A double moved from lane `src` (wave-uniform or not) to the calling lane, as two 32-bit pulls.
__device__ inline double pullD(double v, int src)
{
    const long long x = __double_as_longlong(v);
    const int lo = __shfl((int)x, src, kWave), hi = __shfl((int)(x >> 32), src, kWave);
    return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}

// A double of lane p as a scalar (p wave-uniform).
__device__ inline double laneD(double v, int p)
{
    const long long x = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)x, p), hi = __builtin_amdgcn_readlane((int)(x >> 32), p);
    return __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
}

// Position of the k-th (from 0) set bit of x, which has more than k set bits.
__device__ inline int nthSetBit(uint64_t x, int k)
{
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1)
    {
        const uint64_t lo = x & ((uint64_t(1) << w) - 1);
        const int c = __popcll(lo);
        if (k >= c)
        {
            k -= c;
            x >>= w;
            pos += w;
        }
        else
        {
            x = lo;
        }
    }
    return pos;
}

// Mixed form with shared slots (force_rounds=1, the default): the same lists, culling, tile and pair terms as
// forceKernelMixed, with the accepted pairs of a tile spread evenly over the wavefront.  Lane p
// owns particle p of the chunk (up to 64 at a time).  For each tile:
//   1. the support test of each particle runs over the 64 slots at once (its position read as
//      scalars), giving its mask of accepted slots, kept in the particle's lane;
//   2. with a_p accepted slots for particle p, the smallest m with sum ceil(a_p / m) <= 64 gives
//      each particle ceil(a_p / m) consecutive lanes, each taking up to m of its accepted slots;
//   3. each lane pulls its particle's factors from the owner lane once, evaluates its m pairs
//      into FP32 partials (i- and j-side terms together), and adds them to the particle's
//      partials in LDS (atomic adds within the wavefront, contended only by the lanes sharing
//      the particle);
//   4. the owners add the tile's partials into their FP64 sums.
// A wavefront then runs the pair terms about total / 64 times per tile; with lanes that keep
// their own slots it runs them as often as its busiest lane accepts a pair.  To fit four
// wavefronts per SIMD (128 VGPRs, under 40 KB of LDS per block) the chunk's positions are read
// from their owner lanes, not from LDS, and the chunk's box lives in LDS.
template<bool PER, bool DISC, bool FULL>
__global__ __launch_bounds__(kForceThreads) __attribute__((amdgpu_waves_per_eu(4))) void forceKernelMixedRounds(ForceArgs A)
{
    __shared__ int    segStart[kForceWaves][kWave];
    __shared__ int    segJ[kForceWaves][kWave];
    __shared__ int    tileHead[kForceWaves][kWave];   // staging (tileOwner); in consume, the lanes' particles
    enum { X, Y, Z, VX, VY, VZ, U, ND };
    enum { H, HS, HI, GK, RH, R1, P2, CS, AL, PR, AU, NS };
    __shared__ double td[kForceWaves][ND][kWave];
    __shared__ float  ts[kForceWaves][NS][kWave];
    __shared__ int    jq[kForceWaves][2 * kWave];  // candidates that passed the particle test
    __shared__ double cbox[kForceWaves][8];        // the chunk's box: centre, half-sizes, largest R h
    enum { AX, AY, AZ, AF4, ADIV, AVM, NA };
    __shared__ float  pacc[kForceWaves][NA][kWave];   // the tile's partials per particle (AVM: the max's bits)
    const int w = threadIdx.x >> 6, lane = laneId();
    const size_t S = A.f32Stride;
    auto F = [&](int k, uint32_t idx) { return A.f32[(size_t)k * S + idx]; };
    int a;
    uint32_t p0, p1;
    unsigned long long nAcc = 0, nCand = 0;   // wave-uniform
    if (A.items.resolve(blockIdx.x * kForceWaves + w, A.nLeaves, A.layout, a, p0, p1))
    {
        const int iLeaf = a;
        const uint32_t lo = A.jOffset[a], hi = A.jOffset[a + 1];
        const DBox box = A.boxes[A.leafSet[iLeaf]];
        const float pm = (float)A.pmass, beta = (float)A.beta;
        const float R2 = (float)kRad2;
        for (uint32_t q0 = p0; q0 < p1; q0 += kWave)
        {
            const int c = __builtin_amdgcn_readfirstlane((int)min((uint32_t)kWave, p1 - q0));
            const bool own = lane < c;
            const uint32_t i = q0 + (uint32_t)(own ? lane : 0);
            const double xi = A.x[i], yi = A.y[i], zi = A.z[i], rhi = kRad * A.h[i];
            const double vxi = A.vx[i], vyi = A.vy[i], vzi = A.vz[i], eni = A.u[i];
            const float hif = (float)A.h[i];
            const float hsi = F(0, i), hinvi = F(1, i), gki = F(2, i), rhoi = F(3, i), rho1i = F(4, i);
            const float pro2i = F(5, i), vwi = F(6, i), ali = F(7, i), pri = F(8, i), aui = F(9, i);
            {
                double cbc[3], cbs[3], rChunk;
                waveBox(own, xi, yi, zi, rhi, cbc, cbs, rChunk);
                if (lane == 0)
                {
                    cbox[w][0] = cbc[0]; cbox[w][1] = cbc[1]; cbox[w][2] = cbc[2];
                    cbox[w][3] = cbs[0]; cbox[w][4] = cbs[1]; cbox[w][5] = cbs[2];
                    cbox[w][6] = rChunk;
                }
            }
#pragma unroll
            for (int f = 0; f < NA; ++f) pacc[w][f][lane] = 0.0f;
            waveLdsSync();
            double sx = 0, sy = 0, sz = 0, f4s = 0, divs = 0;
            float vmax = 0.0f;
            auto keep = [&](int jl) {
                double hj = A.hmaxLeaf[jl];
                const double rj = (isfinite(hj) && hj > 0.0) ? kRad * hj : 0.0;
                for (int k = 0; k < c; ++k)
                {
                    const double r = fmax(laneD(rhi, k), rj);
                    if (pointLeafGap2<PER>(laneD(xi, k), laneD(yi, k), laneD(zi, k), A.lbox, A.lboxN, jl, box) <=
                        r * r * kCullSlack)
                        return true;
                }
                return false;
            };
            auto keepJ = [&](int j) {
                const double cc[3] = {cbox[w][0], cbox[w][1], cbox[w][2]};
                const double cs[3] = {cbox[w][3], cbox[w][4], cbox[w][5]};
                const double r = fmax(cbox[w][6], kRad * A.h[j]);
                return pointBoxGap2<PER>(A.x[j], A.y[j], A.z[j], cc, cs, box) <= r * r * kCullSlack;
            };
            auto load = [&](int l, int j) {
                td[w][X][l] = A.x[j]; td[w][Y][l] = A.y[j]; td[w][Z][l] = A.z[j];
                td[w][VX][l] = A.vx[j]; td[w][VY][l] = A.vy[j]; td[w][VZ][l] = A.vz[j]; td[w][U][l] = A.u[j];
                ts[w][H][l] = (float)A.h[j];
                ts[w][HS][l] = F(0, j); ts[w][HI][l] = F(1, j); ts[w][GK][l] = F(2, j); ts[w][RH][l] = F(3, j);
                ts[w][R1][l] = F(4, j); ts[w][P2][l] = F(5, j); ts[w][CS][l] = F(6, j); ts[w][AL][l] = F(7, j);
                ts[w][PR][l] = F(8, j); ts[w][AU][l] = F(9, j);
            };
            auto image = [&](double& dxd, double& dyd, double& dzd) {
                if constexpr (PER)
                {
                    if (fabs(dxd) > 0.5 * box.l[0]) dxd = dxd - box.l[0] * copysign(1.0, dxd);
                    if (fabs(dyd) > 0.5 * box.l[1]) dyd = dyd - box.l[1] * copysign(1.0, dyd);
                    if (fabs(dzd) > 0.5 * box.l[2]) dzd = dzd - box.l[2] * copysign(1.0, dzd);
                }
            };
            auto consume = [&](int nt) {
                nCand += (unsigned long long)nt * (unsigned long long)c;
                // 1. each particle's accepted slots (slot `lane` against particle p), kept in lane p
                uint64_t mine = 0;
                {
                    const bool slot = lane < nt;
                    const int sl = slot ? lane : 0;
                    const double xj = td[w][X][sl], yj = td[w][Y][sl], zj = td[w][Z][sl];
                    const float hj = ts[w][H][sl];
                    const float hj2 = hj * hj;
                    for (int p = 0; p < c; ++p)
                    {
                        // the pair test: 0 < r < R max(h_i, h_j)
                        double dxd = laneD(xi, p) - xj, dyd = laneD(yi, p) - yj, dzd = laneD(zi, p) - zj;
                        image(dxd, dyd, dzd);
                        const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
                        const float dr2 = dx * dx + dy * dy + dz * dz;
                        const float hp =
                            __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, hif), p));
                        const float rmax2 = R2 * fmaxf(hp * hp, hj2) * (1.0f + 1e-6f);
                        const uint64_t m = ballot64(slot && (dr2 > 0.0f) && (dr2 < rmax2));
                        mine = lane == p ? m : mine;
                    }
                }
                // 2. shares: m pairs per lane, ceil(a_p / m) lanes for particle p
                const int ap = __popcll(mine);
                const int total = waveReduceSum(ap);
                if (total == 0) return;
                int m = (total + kWave - 1) / kWave;
                for (;;)
                {
                    const int need = waveReduceSum((ap + m - 1) / m);
                    if (need <= kWave) break;
                    ++m;
                }
                const int lanesP = (ap + m - 1) / m;
                const int startP = waveInclusiveScan(lanesP) - lanesP;
                tileHead[w][lane] = -1;
                if (lanesP > 0) tileHead[w][startP] = lane;
                waveLdsSync();
                const int p = waveInclusiveMax(tileHead[w][lane], -1);   // this lane's particle (-1: none)
                waveLdsSync();
                const int pp = p < 0 ? 0 : p;
                const int r = lane - __shfl(startP, pp, kWave);          // its share of the particle's pairs
                const int apOf = __shfl(ap, pp, kWave);
                const int cnt = p < 0 ? 0 : max(0, min(m, apOf - r * m));
                const uint64_t mk = (uint64_t)(uint32_t)__shfl((int)(uint32_t)mine, pp, kWave) |
                                    ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(mine >> 32), pp, kWave) << 32);
                uint64_t bits = 0;
                if (cnt > 0) bits = mk & ~((uint64_t(1) << nthSetBit(mk, r * m)) - 1);
                // 3. the particle's factors from its owner lane, then the lane's pairs
                const double xp = pullD(xi, pp), yp = pullD(yi, pp), zp = pullD(zi, pp);
                const double vxp = pullD(vxi, pp), vyp = pullD(vyi, pp), vzp = pullD(vzi, pp), enp = pullD(eni, pp);
                const float hp = __shfl(hif, pp, kWave), hsp = __shfl(hsi, pp, kWave);
                const float hinvp = __shfl(hinvi, pp, kWave), gkp = __shfl(gki, pp, kWave);
                const float rhop = __shfl(rhoi, pp, kWave), rho1p = __shfl(rho1i, pp, kWave);
                const float pro2p = __shfl(pro2i, pp, kWave), vwp = __shfl(vwi, pp, kWave);
                const float alp = __shfl(ali, pp, kWave), prp = __shfl(pri, pp, kWave), aup = __shfl(aui, pp, kWave);
                float tx = 0, ty = 0, tz = 0, tf4 = 0, tdiv = 0, tvm = 0;
                for (int t = 0; t < m; ++t)
                {
                    bool counted = false;
                    if (t < cnt)
                    {
                        const int k = __builtin_ctzll(bits);
                        bits &= bits - 1;
                        double dxd = xp - td[w][X][k], dyd = yp - td[w][Y][k], dzd = zp - td[w][Z][k];
                        image(dxd, dyd, dzd);
                        const float dx = (float)dxd, dy = (float)dyd, dz = (float)dzd;
                        const float dr2 = dx * dx + dy * dy + dz * dz;
                        const float hj = ts[w][H][k];
                        const float qi2 = dr2 * hsp, qj2 = dr2 * ts[w][HS][k];
                        const float dr = __builtin_amdgcn_sqrtf(dr2), rij1 = __builtin_amdgcn_rcpf(dr);
                        const float rx = dx * rij1, ry = dy * rij1, rz = dz * rij1;
                        const float dvx = (float)(vxp - td[w][VX][k]), dvy = (float)(vyp - td[w][VY][k]);
                        const float dvz = (float)(vzp - td[w][VZ][k]);
                        const float projv = dvx * rx + dvy * ry + dvz * rz;
                        const float rhoj = ts[w][RH][k], rho1j = ts[w][R1][k], pro2j = ts[w][P2][k];
                        const float vwj = ts[w][CS][k], alj = ts[w][AL][k], prj = ts[w][PR][k], auj = ts[w][AU][k];
                        const float vsigi = fmaxf(vwp - beta * projv, 0.0f);
                        const float vsigavi = fmaxf(alp * vwp - beta * projv, 0.0f);
                        const float vsigj = fmaxf(vwj - beta * projv, 0.0f);
                        const float vsigavj = fmaxf(alj * vwj - beta * projv, 0.0f);
                        const float pvmax = fmaxf(vsigi, vsigj);
                        float qr2i = 0.0f, qr2j = 0.0f;
                        const float denij = (float)(enp - td[w][U][k]);
                        const float vsigu =
                            __builtin_amdgcn_sqrtf(fabsf(prp - prj) * (2.0f * __builtin_amdgcn_rcpf(rhop + rhoj)));
                        if constexpr (DISC)
                        {
                            const float hori = hp * rij1, horj = hj * rij1;
                            if (projv < 0.0f)
                            {
                                qr2i = -0.5f * rho1p * (alp * vwp - beta * projv) * hori * projv;
                                qr2j = -0.5f * rho1j * (alj * vwj - beta * projv) * horj * projv;
                            }
                            else
                            {
                                qr2i = -0.5f * rho1p * alp * vwp * hori * projv;
                                qr2j = -0.5f * rho1j * alj * vwj * horj * projv;
                            }
                        }
                        else if (projv < 0.0f)
                        {
                            qr2i = -0.5f * rho1p * vsigavi * projv;
                            qr2j = -0.5f * rho1j * vsigavj * projv;
                        }
                        if (qi2 < R2)
                        {
                            counted = true;
                            tvm = fmaxf(tvm, pvmax);
                            float wk, dwk;
                            const float q = dr * hinvp;
                            kernWdW<float>(qi2, q, wk, dwk);
                            tdiv += pm * dwk * projv;
                            const float gk = dwk * gkp;
                            const float gp = pm * (pro2p + qr2i) * gk;
                            tx -= gp * rx; ty -= gp * ry; tz -= gp * rz;
                            const float pdvt = pm * pro2p * projv * gk;
                            const float diss = DISC ? -0.5f * pm * rho1p * alp * vwp * (hp * rij1) * projv * projv * gk
                                                    : pm * qr2i * projv * gk;
                            const float dend = vsigu * denij * aup * gk;
                            if constexpr (FULL) { tf4 += pdvt; tf4 += diss; }
                            else
                            {
                                if (A.pdv) tf4 += pdvt;
                                if (A.shock) tf4 += diss;
                            }
                            tf4 += dend;
                        }
                        if (qj2 < R2)
                        {
                            counted = true;
                            tvm = fmaxf(tvm, pvmax);
                            float wk, dwk;
                            const float q = dr * ts[w][HI][k];
                            kernWdW<float>(qj2, q, wk, dwk);
                            const float gk = dwk * ts[w][GK][k];
                            const float gp = pm * (pro2j + qr2j) * gk;
                            tx -= gp * rx; ty -= gp * ry; tz -= gp * rz;
                            tf4 += vsigu * denij * auj * gk;
                        }
                    }
                    nAcc += (unsigned long long)__popcll(ballot64(counted));
                }
                // 4. the lane's partials into its particle's
                if (cnt > 0)
                {
                    atomicAdd(&pacc[w][AX][pp], tx); atomicAdd(&pacc[w][AY][pp], ty);
                    atomicAdd(&pacc[w][AZ][pp], tz); atomicAdd(&pacc[w][AF4][pp], tf4);
                    atomicAdd(&pacc[w][ADIV][pp], tdiv);
                    atomicMax(reinterpret_cast<unsigned*>(&pacc[w][AVM][pp]), __float_as_uint(tvm) & 0x7FFFFFFFu);
                }
                waveLdsSync();
                if (own)
                {
                    sx += (double)pacc[w][AX][lane]; sy += (double)pacc[w][AY][lane]; sz += (double)pacc[w][AZ][lane];
                    f4s += (double)pacc[w][AF4][lane]; divs += (double)pacc[w][ADIV][lane];
                    vmax = fmaxf(vmax, pacc[w][AVM][lane]);
                }
#pragma unroll
                for (int f = 0; f < NA; ++f) pacc[w][f][lane] = 0.0f;
                waveLdsSync();
            };
            forEachTileCompact(segStart[w], segJ[w], tileHead[w], jq[w], A.layout, A.jList, lo, hi, keep, keepJ, load,
                               consume);
            if (own)
            {
                A.fx[i] = sx; A.fy[i] = sy; A.fz[i] = sz;
                A.f4[i] = f4s; A.vsig[i] = (double)vmax;
                A.divv[i] = -divs * A.divfac[i];
            }
            waveLdsSync();   // cbox is rewritten for the next chunk
        }
    }
    if (laneId() == 0 && (nAcc | nCand))
    {
        unsigned long long* slot = A.counters + ((blockIdx.x * kForceWaves + w) % kCounterSlots) * kCounterStride;
        atomicAdd(slot + 0, nAcc);
        atomicAdd(slot + 1, nCand);
    }
}

// FP32 copies of the force factors (ForceArgs::f32 layout).
__global__ void forceF32Kernel(ForceArgs A, int n, float* __restrict__ dst)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t S = A.f32Stride;
    const double* src[10] = {A.hsqinv, A.hinv, A.grkfac, A.rhoh, A.rho1, A.pro2, A.cs, A.alpha, A.pres, A.auterm};
#pragma unroll
    for (int k = 0; k < 10; ++k) dst[(size_t)k * S + i] = (float)src[k][i];
}

__global__ void deadForceKernel(int first, int n, double* __restrict__ fx, double* __restrict__ fy,
                                double* __restrict__ fz, double* __restrict__ f4, double* __restrict__ vsig,
                                double* __restrict__ divv)
{
    const int i = first + blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    fx[i] = 0.0; fy[i] = 0.0; fz[i] = 0.0; f4[i] = 0.0; vsig[i] = 0.0; divv[i] = 0.0;
}

void launchForcePrep(const double* h, const double* gradh, const double* pro2, double* hsqinv, double* hinv,
                     double* rhoh, double* rho1, double* grkfac, double* pres, double* auterm, double* divfac, int n,
                     double pmass, double alphau, double hfact, hipStream_t st)
{
    if (n <= 0) return;
    forcePrepKernel<<<iceil(n, 256), 256, 0, st>>>(h, gradh, pro2, hsqinv, hinv, rhoh, rho1, grkfac, pres, auterm,
                                                   divfac, n, pmass, alphau, hfact);
    LAUNCH_CHECK();
}

void launchForceF32(const ForceArgs& a, int n, float* dst, hipStream_t st)
{
    if (n <= 0) return;
    forceF32Kernel<<<iceil(n, 256), 256, 0, st>>>(a, n, dst);
    LAUNCH_CHECK();
}

void launchForce(const ForceArgs& a, bool periodic, bool disc, bool full, bool mixed, bool rounds, hipStream_t st)
{
    if (a.nLeaves <= 0) return;
    const int grid = iceil(a.items.count(a.nLeaves), kForceWaves);
#define CSPH_F(P, D, F)                                                                           \
    do {                                                                                          \
        if (mixed && rounds) forceKernelMixedRounds<P, D, F><<<grid, kForceThreads, 0, st>>>(a);  \
        else if (mixed)      forceKernelMixed<P, D, F><<<grid, kForceThreads, 0, st>>>(a);        \
        else                 forceKernel<P, D, F><<<grid, kForceThreads, 0, st>>>(a);             \
    } while (0)
    if (periodic)
    {
        if (disc) { if (full) CSPH_F(true, true, true); else CSPH_F(true, true, false); }
        else      { if (full) CSPH_F(true, false, true); else CSPH_F(true, false, false); }
    }
    else
    {
        if (disc) { if (full) CSPH_F(false, true, true); else CSPH_F(false, true, false); }
        else      { if (full) CSPH_F(false, false, true); else CSPH_F(false, false, false); }
    }
#undef CSPH_F
    LAUNCH_CHECK();
}

void launchDeadForce(int first, int n, double* fx, double* fy, double* fz, double* f4, double* vsig, double* divv,
                     hipStream_t st)
{
    if (n > first)
    {
        deadForceKernel<<<iceil(n - first, 256), 256, 0, st>>>(first, n, fx, fy, fz, f4, vsig, divv);
        LAUNCH_CHECK();
    }
}

} // namespace csph
