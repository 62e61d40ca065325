// DEM kernels of the coupled step on gfx950, all FP64 (compiled without contraction so the
// arithmetic rounds like the reference's scalar code):
//   * neighbour list by cell sort (DEM::neighborList, DEM.cpp:473-652): counting sort of
//     particles into link cells, per-particle gather over the 27 neighbour cells, partners
//     sorted ascending, tangential history carried over by partner id
//   * forces (DEM::evaluateForces, DEM.cpp:655-796): one thread per particle evaluates every
//     pair from its side with the reference's (n1 > n2) orientation, so each pair's force is
//     computed bit-identically twice and no atomics are needed; the pair's tangential history
//     is read from the previous buffer and written once, by the thread of the higher index
//   * integration and periodic wrap (DEM.cpp:2676-2836)
#include "lbdem_internal.hpp"
#include "launch.hpp"

#include <algorithm>
#include <climits>

namespace lbdem {

__device__ inline double dmaxd(double a, double b) { return a > b ? a : b; }
__device__ inline int n_int(double x) { return x >= 0 ? int(x + 0.5) : int(x - 0.5); }
__device__ inline int s_int(double x) { return x >= 0 ? int(x) : int(x - 1.0); }
__device__ inline double dbox(double a, double b) { return a - s_int(a / b) * b; }
__device__ inline double n_image(double a, double b) { return a - n_int(a / b) * b; }

struct V { double x, y, z; };
__device__ inline V ld3(const double* p, int i) { return V{p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__device__ inline void st3(double* p, int i, V v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
__device__ inline V add(V a, V b) { return V{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V sub(V a, V b) { return V{a.x - b.x, a.y - b.y, a.z - b.z}; }

// ------------------------------------------------------------------------------------------
// neighbour-list trigger (DEM.cpp:366-378, 460-470)
// ------------------------------------------------------------------------------------------
// The trigger needs the largest |v|^2 over the particles at the start of the step: the
// velocities the previous step's k_dem_integrate wrote, which folds them into the VEL_SLOTS slots
// of velMax with atomic maxima on the bit patterns (block b into slot b % VEL_SLOTS, so no slot
// takes more than a few atomics; non-negative doubles order like their bit patterns; zero at
// initialisation, when every velocity is zero). k_maxdisp (one wavefront) reads and clears them.
static_assert(VEL_SLOTS == 64, "k_maxdisp reduces the slots in one wavefront");
__device__ inline void k_maxdisp(unsigned long long* velMax, double deltat, DevScalars* sc, int force) {
    const int t = threadIdx.x;
    double m = 0.0;
    if (t < VEL_SLOTS) {
        m = __longlong_as_double(velMax[t]);
        velMax[t] = 0ull;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if (t != 0) return;
    if (force) {
        sc->rebuild = 1;
        sc->builds += 1;
        return;
    }
    const double maxVel = sqrt(m);
    sc->maxDisp += maxVel * deltat;
    if (sc->maxDisp > 0.05) {
        sc->maxDisp = 0.0;
        sc->rebuild = 1;
        sc->newNeighborList = 1;
        sc->builds += 1;
    } else {
        sc->rebuild = 0;
    }
}

// ------------------------------------------------------------------------------------------
// cell list (counting sort by link cell)
// ------------------------------------------------------------------------------------------
struct NbrArgs {
    DemParams p;
    const double* x0;
    const double* r;
    int* cellCount;
    int* cellStart;
    int* cellFill;
    int* cellItems;
    int* cellKey;
    int* nbrCount;
    int* nbrList;
    int* nbrSlot;
    DevScalars* sc;
};

__device__ inline void k_cells_assign(NbrArgs a) {
    if (!a.sc->rebuild) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.p.nPart) return;
    const int lx = int(dbox(a.x0[3 * p] / a.p.demSize[0], 1) * a.p.nlx);
    const int ly = int(dbox(a.x0[3 * p + 1] / a.p.demSize[1], 1) * a.p.nly);
    const int lz = int(dbox(a.x0[3 * p + 2] / a.p.demSize[2], 1) * a.p.nlz);
    int key = lx * a.p.nly * a.p.nlz + ly * a.p.nlz + lz;
    const int ncell = a.p.nlx * a.p.nly * a.p.nlz;
    key = key < 0 ? 0 : (key >= ncell ? ncell - 1 : key);
    a.cellKey[p] = key;
    atomicAdd(&a.cellCount[key], 1);
}
// exclusive prefix sum of the cell counts by one block of 1024 threads, 4096 cells at a time: each
// thread takes 4 consecutive counts (coalesced over the block), the block scans the thread sums
// (wavefront shuffles, wavefront totals in LDS) and the chunk's total carries into the next chunk
constexpr int SCAN_THREADS = 1024, SCAN_PER = 4;
__device__ inline void k_cells_scan(NbrArgs a, int ncell) {
    if (!a.sc->rebuild) return;
    __shared__ int wtot[SCAN_THREADS / 64];
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int c0 = 0; c0 < ncell; c0 += SCAN_THREADS * SCAN_PER) {
        const int b = c0 + SCAN_PER * int(threadIdx.x);
        int v[SCAN_PER];
        int s = 0;
#pragma unroll
        for (int j = 0; j < SCAN_PER; ++j) {
            v[j] = b + j < ncell ? a.cellCount[b + j] : 0;
            s += v[j];
        }
        int incl = s;  // inclusive scan over the wavefront
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int o = __shfl_up(incl, off, 64);
            if (lane >= off) incl += o;
        }
        if (lane == 63) wtot[wv] = incl;
        __syncthreads();
        int acc = carry + incl - s;
        for (int w = 0; w < wv; ++w) acc += wtot[w];
#pragma unroll
        for (int j = 0; j < SCAN_PER; ++j) {
            if (b + j < ncell) a.cellStart[b + j] = acc;
            acc += v[j];
        }
        __syncthreads();  // carry and wtot are read
        if (threadIdx.x == SCAN_THREADS - 1) carry = acc;
        __syncthreads();
    }
}
__device__ inline void k_cells_fill(NbrArgs a) {
    if (!a.sc->rebuild) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.p.nPart) return;
    const int key = a.cellKey[p];
    const int pos = atomicAdd(&a.cellFill[key], 1);
    a.cellItems[a.cellStart[key] + pos] = p;
}

// the pair (n1, n2) = (higher, lower index) within sig + 2 (DEM.cpp:601-608); positions and radii
// passed in, differences taken higher minus lower as in the reference
__device__ inline bool within_cut(const DemParams& P, const double* x1, double r1, const double* x2, double r2) {
#pragma clang fp contract(off)
    const double x12 = n_image(x1[0] - x2[0], P.demSize[0]);
    const double y12 = n_image(x1[1] - x2[1], P.demSize[1]);
    const double z12 = n_image(x1[2] - x2[2], P.demSize[2]);
    const double r12 = x12 * x12 + y12 * y12 + z12 * z12;
    const double sig = r1 + r2;
    const double cut = sig + 2.0;
    return r12 <= cut * cut;
}

// Partner gather (pairs with minimum-image distance within sig + 2 in all axes), NBR_LANES lanes
// per particle: with link cells, lane l < 27 takes the wrapped neighbour cell offset
// (l / 9 - 1, l / 3 % 3 - 1, l % 3 - 1) (a cell that two offsets reach, with fewer than 3 cells
// along z, is walked by both, so its partners are listed twice, as the reference lists them); in
// the direct sum lane l takes every NBR_LANES-th particle. Partners collect in LDS; each lane then
// ranks its entries among the group's (ascending, equal ids by collection order) and writes them
// to their place in the list.
constexpr int NBR_LANES = 32, NBR_BLOCK = 256, NBR_GROUPS = NBR_BLOCK / NBR_LANES;
__device__ inline void k_nbr_build(NbrArgs a) {
    if (!a.sc->rebuild) return;
    __shared__ int sl[NBR_GROUPS][NBR_CAP];
    __shared__ int sc[NBR_GROUPS];
    __shared__ int sover[NBR_GROUPS];
    const int grp = threadIdx.x / NBR_LANES, l = threadIdx.x % NBR_LANES;
    const int p = blockIdx.x * NBR_GROUPS + grp;
    const bool valid = p < a.p.nPart;
    if (l == 0) { sc[grp] = 0; sover[grp] = 0; }
    __syncthreads();
    double xp[3] = {0.0, 0.0, 0.0}, rp = 0.0;
    if (valid) { xp[0] = a.x0[3 * p]; xp[1] = a.x0[3 * p + 1]; xp[2] = a.x0[3 * p + 2]; rp = a.r[p]; }
    auto consider = [&](int q) {
        if (q == p) return;
        const double xq[3] = {a.x0[3 * q], a.x0[3 * q + 1], a.x0[3 * q + 2]};
        const double rq = a.r[q];
        if (!(p > q ? within_cut(a.p, xp, rp, xq, rq) : within_cut(a.p, xq, rq, xp, rp))) return;
        const int k = atomicAdd(&sc[grp], 1);
        if (k < NBR_CAP - 1) sl[grp][k] = q; else sover[grp] = 1;  // 63 partners at most
    };
    if (valid) {
        if (a.p.nlx == 0) {
            for (int q = l; q < a.p.nPart; q += NBR_LANES) consider(q);
        } else if (a.p.nlz > 0 && l < 27) {
            // with nlz < 3 the wrapped z offsets repeat a cell, and the partners in it are taken once
            // per visit (DEM.cpp:585-627); with nlz == 0 the reference builds no list
            const int nlx = a.p.nlx, nly = a.p.nly, nlz = a.p.nlz;
            const int key = a.cellKey[p];
            const int lx = key / (nly * nlz), ly = (key / nlz) % nly, lz = key % nlz;
            const int dx = l / 9 - 1, dy = (l / 3) % 3 - 1, dz = l % 3 - 1;
            const int c = ((lx + dx + nlx) % nlx) * nly * nlz + ((ly + dy + nly) % nly) * nlz + (lz + dz + nlz) % nlz;
            const int b = a.cellStart[c], e = b + a.cellCount[c];
            for (int t = b; t < e; ++t) consider(a.cellItems[t]);
        }
    }
    __syncthreads();
    if (!valid) return;
    const int cnt = sc[grp] < NBR_CAP - 1 ? sc[grp] : NBR_CAP - 1;
    const int* v = sl[grp];
    for (int k = l; k < cnt; k += NBR_LANES) {
        const int x = v[k];
        int rank = 0;
        for (int j = 0; j < cnt; ++j) {
            const int y = v[j];
            rank += (y < x || (y == x && j < k)) ? 1 : 0;
        }
        a.nbrList[p * NBR_CAP + rank] = x;
    }
    if (l == 0) {
        a.nbrCount[p] = cnt;
        if (sover[grp]) atomicOr(&a.sc->nbrOverflow, 1);
    }
}

// Slot and carried tangential history of every entry, SLOT_LANES lanes per particle (a
// particle's lanes lie in one wavefront): the pair's history lives in the entry of the higher
// index, carried over by partner id from the old list (DEM.cpp:633-651; the last matching old
// entry wins, as in the reference's loop); a partner absent from the old list starts with zero
// history. The lower index finds its entry in the partner's (ascending) list. Every lane first
// reads what it needs of its particle's old list and history, then the lanes write the new list,
// history and slots over the old ones in place: a particle's old entries are read and written only
// by its own lanes, and the memory operations of a wavefront are performed in order.
constexpr int SLOT_LANES = 16, SLOT_PER = NBR_CAP / SLOT_LANES;
static_assert(64 % SLOT_LANES == 0, "a particle's lanes lie in one wavefront");
__device__ inline void k_nbr_slots(NbrArgs a, double* xi, int* oldCount, int* oldList) {
    if (!a.sc->rebuild) return;
    // k_nbr_build was the cell list's last reader: clear it for the next rebuild (zero at
    // initialisation)
    const int ncell = a.p.nlx * a.p.nly * a.p.nlz;
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncell; c += gridDim.x * blockDim.x) {
        a.cellCount[c] = 0;
        a.cellFill[c] = 0;
    }
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = e / SLOT_LANES, l = e % SLOT_LANES;
    const bool valid = p < a.p.nPart;
    const int cnt = valid ? a.nbrCount[p] : 0;
    const int oc = valid ? oldCount[p] : 0;
    int q[SLOT_PER], slot[SLOT_PER];
    double h[SLOT_PER][3];
#pragma unroll
    for (int i = 0; i < SLOT_PER; ++i) {
        const int k = l + SLOT_LANES * i;
        q[i] = -1;
        slot[i] = -1;
        h[i][0] = h[i][1] = h[i][2] = 0.0;
        if (k >= cnt) continue;
        q[i] = a.nbrList[p * NBR_CAP + k];
        if (p > q[i]) {
            slot[i] = p * NBR_CAP + k;
            int last = -1;
            for (int m = 0; m < oc; ++m)
                if (oldList[p * NBR_CAP + m] == q[i]) last = m;
            if (last >= 0) {
                const int os = p * NBR_CAP + last;
                h[i][0] = xi[3 * os]; h[i][1] = xi[3 * os + 1]; h[i][2] = xi[3 * os + 2];
            }
        } else {
            const int qq = q[i], qc = a.nbrCount[qq];
            int lo = 0, hi = qc - 1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                const int v = a.nbrList[qq * NBR_CAP + mid];
                if (v == p) { slot[i] = qq * NBR_CAP + mid; break; }
                if (v < p) lo = mid + 1; else hi = mid - 1;
            }
        }
    }
    // every lane's reads of the old entries are issued before any lane overwrites them
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (!valid) return;
    if (l == 0) oldCount[p] = cnt;
#pragma unroll
    for (int i = 0; i < SLOT_PER; ++i) {
        const int k = l + SLOT_LANES * i;
        if (k >= cnt) continue;
        const int s = p * NBR_CAP + k;
        a.nbrSlot[s] = slot[i];
        oldList[s] = q[i];
        if (p > q[i]) { xi[3 * s] = h[i][0]; xi[3 * s + 1] = h[i][1]; xi[3 * s + 2] = h[i][2]; }
    }
}

// ------------------------------------------------------------------------------------------
// forces
// ------------------------------------------------------------------------------------------
struct ForceArgs {
    DemParams p;
    const double *x0, *x1, *w1, *r;
    double *FContact, *MContact, *FLub, *MLub, *FRep;
    double *xiw, *xio;
    const int* nbrCount;
    const int* nbrList;
    const int* nbrSlot;
    const double* xiOld;
    double* xiNew;
    const double *ox0, *ox1, *ow1, *or_;
    double* objContrib;   // per particle, per object: Fc(3) Mc(3) Fl(3) Ml(3)
    double* wallContrib;  // per particle, per wall: Fc(3) Fl(3)
    int demTimeStep;
    double* stressPart;   // SP_REC per particle (with the stress diagnostics)
    const double* shydro; // 9 per particle
};

// S1, S2: the two spheres' stresslets as xx yy zz yz zx xy (with STRESS only)
// normal: within the normal range (the pair counts in lub_coord, DEM.cpp:2360-2363)
struct LubOut { V F, T1, T2; bool act; double S1[6], S2[6]; bool normal; };

// pair lubrication (DEM.cpp:1140-1340, 2302-2521) for the ordered pair (a = n1, b = n2)
template <bool STRESS>
__device__ LubOut pair_lub(const DemParams& P, V xa, V xb, double ra, double rb, V va, V vb, V wa, V wb, double bnum,
                           double bden) {
#pragma clang fp contract(off)
    LubOut o{};
    const double lub_N = P.lubNormal, lub_T = P.lubTangential, lub_R = P.lubRotation, cut = P.lubCutoff;
    const double rad = 0.5 * (ra + rb);
    double x12 = xa.x - xb.x, y12 = xa.y - xb.y, z12 = xa.z - xb.z;
    const double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
    const double umx = va.x - vb.x, umy = va.y - vb.y, umz = va.z - vb.z;
    const double dr = r12 - 2.0 * rad;
    const double h = dmaxd(dr, cut);
    o.act = dr < dmaxd(cut, dmaxd(lub_N, dmaxd(lub_T, lub_R)));
    if (!o.act) return o;
    double Xa, Ya, Yb1, Yb2, Yc11, Yc12, Yc21, Yc22, h_c, beta;
    double Xg1 = 0, Xg2 = 0, Yg1 = 0, Yg2 = 0, Yh11 = 0, Yh12 = 0, Yh21 = 0, Yh22 = 0;  // stresslet (DEM.cpp:2362-2406)
    const double alpha = M_PI * P.kinVisc * rad * P.deltat;
    const double beta1 = bnum / bden;
    const double beta2 = beta1 * beta1, beta3 = beta1 * beta2;
    const double beta4 = pow((1 + beta1), 4.0);
    const double beta5 = STRESS ? pow((1 + beta1), 5.0) : 1.0;
    if (dr < dmaxd(cut, lub_N)) {
        o.normal = true;
        Xa = 4.0 * beta2 / beta4;
        h_c = dmaxd(dr, lub_N);
        Xa *= rad / h - rad / h_c;
        Xa *= 6.0 * alpha;
        if constexpr (STRESS) {
            Xg1 = 12.0 * beta2 / beta5; Xg2 = -12.0 * beta3 / beta5;
            Xg1 *= rad / h - rad / h_c; Xg2 *= rad / h - rad / h_c;
            Xg1 *= 4.0 * alpha * rad; Xg2 *= 4.0 * alpha * rad;
        }
    } else Xa = 0.0;
    if (dr < dmaxd(cut, lub_T)) {
        Ya = 8.0 * beta1 * (2.0 + beta1 + 2.0 * beta2) / (15.0 * beta4);
        Yb1 = -4.0 * beta1 * (4.0 + beta1) / (5.0 * beta4);
        Yb2 = 4.0 * beta1 * (4.0 * beta2 + beta1) / (5.0 * beta4);
        h_c = dmaxd(dr, lub_T);
        beta = alpha * log(h_c / h);
        Ya *= 6.0 * beta; Yb1 *= 4.0 * beta * rad; Yb2 *= 4.0 * beta * rad;
        if constexpr (STRESS) {
            Yg1 = 2.0 * beta1 * (4.0 - beta1 + 7.0 * beta2) / (5.0 * beta5);
            Yg2 = -2.0 * beta2 * (4.0 * beta2 - beta1 + 7.0) / (5.0 * beta5);
            Yg1 *= 4.0 * beta * rad; Yg2 *= 4.0 * beta * rad;
        }
    } else Ya = Yb1 = Yb2 = 0.0;
    if (dr < dmaxd(cut, lub_R)) {
        Yc11 = 16.0 * beta1 / (5.0 * beta4);
        Yc12 = 4.0 * beta2 / (5.0 * beta4);
        Yc21 = 4.0 * beta2 / (5.0 * beta4);
        Yc22 = 16.0 * beta3 / (5.0 * beta4);
        h_c = dmaxd(dr, lub_R);
        beta = alpha * log(h_c / h);
        Yc11 *= 8.0 * beta * rad * rad; Yc12 *= 8.0 * beta * rad * rad;
        Yc21 *= 8.0 * beta * rad * rad; Yc22 *= 8.0 * beta * rad * rad;
        if constexpr (STRESS) {
            Yh11 = 4.0 * beta1 * (2.0 - beta1) / (5.0 * beta5);
            Yh12 = 2.0 * beta2 * (1.0 + 7.0 * beta1) / (5.0 * beta5);
            Yh21 = 2.0 * beta2 * (beta1 + 7.0) / (5.0 * beta5);
            Yh22 = 4.0 * beta2 * (2.0 * beta2 - beta1) / (5.0 * beta5);
            Yh11 *= 8.0 * beta * rad * rad; Yh21 *= 8.0 * beta * rad * rad;
            Yh12 *= 8.0 * beta * rad * rad; Yh22 *= 8.0 * beta * rad * rad;
        }
    } else Yc11 = Yc12 = Yc21 = Yc22 = 0.0;
    x12 /= r12; y12 /= r12; z12 /= r12;
    const double w1x = wa.x, w1y = wa.y, w1z = wa.z, w2x = wb.x, w2y = wb.y, w2z = wb.z;
    const double um_r = umx * x12 + umy * y12 + umz * z12;
    const double w1_r = w1x * x12 + w1y * y12 + w1z * z12;
    const double w2_r = w2x * x12 + w2y * y12 + w2z * z12;
    const double umxrx = umy * z12 - umz * y12, umxry = umz * x12 - umx * z12, umxrz = umx * y12 - umy * x12;
    const double w1xrx = w1y * z12 - w1z * y12, w1xry = w1z * x12 - w1x * z12, w1xrz = w1x * y12 - w1y * x12;
    const double w2xrx = w2y * z12 - w2z * y12, w2xry = w2z * x12 - w2x * z12, w2xrz = w2x * y12 - w2y * x12;
    double Fx = -Xa * um_r * x12 - Ya * (umx - um_r * x12);
    double Fy = -Xa * um_r * y12 - Ya * (umy - um_r * y12);
    double Fz = -Xa * um_r * z12 - Ya * (umz - um_r * z12);
    Fx += -Yb1 * w1xrx + Yb2 * w2xrx;
    Fy += -Yb1 * w1xry + Yb2 * w2xry;
    Fz += -Yb1 * w1xrz + Yb2 * w2xrz;
    double T1x = Yb1 * umxrx, T1y = Yb1 * umxry, T1z = Yb1 * umxrz;
    double T2x = -Yb2 * umxrx, T2y = -Yb2 * umxry, T2z = -Yb2 * umxrz;
    T1x += -Yc11 * (w1x - w1_r * x12) - Yc12 * (w2x - w2_r * x12);
    T1y += -Yc11 * (w1y - w1_r * y12) - Yc12 * (w2y - w2_r * y12);
    T1z += -Yc11 * (w1z - w1_r * z12) - Yc12 * (w2z - w2_r * z12);
    T2x += -Yc21 * (w1x - w1_r * x12) - Yc22 * (w2x - w2_r * x12);
    T2y += -Yc21 * (w1y - w1_r * y12) - Yc22 * (w2y - w2_r * y12);
    T2z += -Yc21 * (w1z - w1_r * z12) - Yc22 * (w2z - w2_r * z12);
    o.F = V{Fx, Fy, Fz}; o.T1 = V{T1x, T1y, T1z}; o.T2 = V{T2x, T2y, T2z};
    if constexpr (STRESS) {  // DEM.cpp:2454-2477
        const double xs[3] = {x12, y12, z12}, us[3] = {umx, umy, umz};
        const double w1s[3] = {w1xrx, w1xry, w1xrz}, w2s[3] = {w2xrx, w2xry, w2xrz};
        // components xx yy zz yz zx xy as (i, j) pairs
        const int ci[6] = {0, 1, 2, 1, 2, 0}, cj[6] = {0, 1, 2, 2, 0, 1};
#pragma unroll
        for (int e = 0; e < 6; ++e) {
            const int i = ci[e], j = cj[e];
            double s1 = Xg1 * um_r * xs[i] * xs[j] + Yg1 * (xs[i] * us[j] + us[i] * xs[j]);
            double s2 = -Xg2 * um_r * xs[i] * xs[j] - Yg2 * (xs[i] * us[j] + us[i] * xs[j]);
            s1 += -Yh11 * (xs[i] * w1s[j] + w1s[i] * xs[j]) - Yh12 * (xs[i] * w2s[j] + w2s[i] * xs[j]);
            s2 += -Yh21 * (xs[i] * w1s[j] + w1s[i] * xs[j]) - Yh22 * (xs[i] * w2s[j] + w2s[i] * xs[j]);
            o.S1[e] = s1; o.S2[e] = s2;
        }
    }
    return o;
}

// S += -a (x) f, row-major (the contact and electrostatic stresses, DEM.cpp:1715-1783, 1451-1479)
__device__ inline void add_arm(double* S, double ax, double ay, double az, double fx, double fy, double fz) {
#pragma clang fp contract(off)
    S[0] += -ax * fx; S[1] += -ax * fy; S[2] += -ax * fz;
    S[3] += -ay * fx; S[4] += -ay * fy; S[5] += -ay * fz;
    S[6] += -az * fx; S[7] += -az * fy; S[8] += -az * fz;
}

// wall-type contact with the shared per-particle history xiw (DEM.cpp:1946-2153, 2157-2296)
__device__ void wall_like_contact(const DemParams& P, double dampCoeff, double r, V x1, V w1, V& xi, double x1w,
                                  double y1w, double z1w, V& Fout, V& Mout, V* felOut = nullptr,
                                  V* fvisOut = nullptr) {
#pragma clang fp contract(off)
    V fel{0, 0, 0}, fvis{0, 0, 0};  // elastic and viscous parts (DEM.cpp:2001-2079)
    const double r1w = sqrt(x1w * x1w + y1w * y1w + z1w * z1w);
    const double delw = r - r1w;
    double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0, n1wx = 0, n1wy = 0, n1wz = 0;
    if (delw > 0.0) {
        n1wx = x1w / r1w; n1wy = y1w / r1w; n1wz = z1w / r1w;
        const double gn = x1.x * n1wx + x1.y * n1wy + x1.z * n1wz;
        const double k = P.linearStiff, mu = P.frictionCoefPart;
        fnx = k * delw * n1wx - dampCoeff * gn * n1wx;
        fny = k * delw * n1wy - dampCoeff * gn * n1wy;
        fnz = k * delw * n1wz - dampCoeff * gn * n1wz;
        fel = V{k * delw * n1wx, k * delw * n1wy, k * delw * n1wz};
        fvis = V{-dampCoeff * gn * n1wx, -dampCoeff * gn * n1wy, -dampCoeff * gn * n1wz};
        if (mu != 0.0) {
            const double ox = r * w1.x, oy = r * w1.y, oz = r * w1.z;
            const double vx = x1.x + n1wy * oz - n1wz * oy - gn * n1wx;
            const double vy = x1.y + n1wz * ox - n1wx * oz - gn * n1wy;
            const double vz = x1.z + n1wx * oy - n1wy * ox - gn * n1wz;
            const double xin = xi.x * n1wx + xi.y * n1wy + xi.z * n1wz;
            xi.x -= xin * n1wx; xi.y -= xin * n1wy; xi.z -= xin * n1wz;
            const double kt = 0.2 * k;
            const double dct = P.kinVisc == 0.0 ? dampCoeff : 0.0;
            ftx = -kt * xi.x - dct * vx; fty = -kt * xi.y - dct * vy; ftz = -kt * xi.z - dct * vz;
            const double fnn = sqrt(fnx * fnx + fny * fny + fnz * fnz);
            const double ftt = sqrt(ftx * ftx + fty * fty + ftz * ftz);
            const double fri = mu * fnn;
            const double tx = ftx / ftt, ty = fty / ftt, tz = ftz / ftt;
            xi.x += vx * P.deltat; xi.y += vy * P.deltat; xi.z += vz * P.deltat;
            if (fri < ftt) {
                ftx = fri * tx; fty = fri * ty; ftz = fri * tz;
                xi.x = -(1 / kt) * (ftx + dct * vx);
                xi.y = -(1 / kt) * (fty + dct * vy);
                xi.z = -(1 / kt) * (ftz + dct * vz);
            }
            fel = add(fel, V{ftx, fty, ftz});
        }
    } else {
        xi = V{0, 0, 0};
    }
    if (felOut) *felOut = fel;
    if (fvisOut) *fvisOut = fvis;
    Fout = V{fnx + ftx, fny + fty, fnz + ftz};
    const double dtx = fty * n1wz - ftz * n1wy, dty = ftz * n1wx - ftx * n1wz, dtz = ftx * n1wy - fty * n1wx;
    Mout = V{r * dtx, r * dty, r * dtz};
}

constexpr int DEM_LANES = 4;  // lanes per particle in the force kernel (all wavefronts resident at once on the 31 000-sphere run)

// sum over the DEM_LANES lanes of a particle's group, in a fixed butterfly order
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int m = 1; m < DEM_LANES; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ inline V group_sum(V v) { return V{group_sum(v.x), group_sum(v.y), group_sum(v.z)}; }

// Forces on particle p: DEM_LANES lanes per particle share its Verlet partners (lane l takes
// partners l, l + DEM_LANES, ...), their partial sums are combined in a fixed order, and the
// group's first lane adds the wall and impactor terms and writes the results. STRESS also
// accumulates the particle's stress tensors (DEM.cpp:655-771) into its SP_REC record.
template <bool STRESS>
__device__ inline void k_dem_forces(ForceArgs a, double dampCoeff) {
#pragma clang fp contract(off)
    const int gt = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = gt % DEM_LANES;
    const DemParams& P = a.p;
    const bool valid = gt / DEM_LANES < P.nPart;
    const int p = valid ? gt / DEM_LANES : 0;  // every lane takes part in the group sums
    const V xp = ld3(a.x0, p), vp = ld3(a.x1, p), wp = ld3(a.w1, p);
    const double rp = a.r[p];
    V Fc{0, 0, 0}, Mc{0, 0, 0}, Fl{0, 0, 0}, Ml{0, 0, 0}, Fr{0, 0, 0};
    double Sl[6] = {0, 0, 0, 0, 0, 0};  // SLub as xx yy zz yz zx xy
    double Sc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Sr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double Se[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Sv[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // SElas, SVisc
    V Fe{0, 0, 0};
    double coord = 0.0, lubc = 0.0;
    const int cnt = valid ? a.nbrCount[p] : 0;
    const bool waited = a.demTimeStep >= P.objWaitTime;
    // One pass over the lane's partners: the pair's data are gathered once and the three pair
    // terms follow in the reference's order; each accumulator sees its terms in partner order, as
    // three separate loops would give them.
    for (int k = lane; k < cnt; k += DEM_LANES) {
        const int q = a.nbrList[p * NBR_CAP + k];
        const int n1 = p > q ? p : q, n2 = p > q ? q : p;
        const int slot = a.nbrSlot[p * NBR_CAP + k];
        const V A = ld3(a.x0, n1), B = ld3(a.x0, n2), vA = ld3(a.x1, n1), vB = ld3(a.x1, n2);
        const V wA = ld3(a.w1, n1), wB = ld3(a.w1, n2);
        const double rA = a.r[n1], rB = a.r[n2];
        V xi{0, 0, 0};
        if (slot >= 0) xi = ld3(a.xiOld, slot);
        // particle lubrication (DEM.cpp:2302-2521)
        if (waited) {
            const LubOut o = pair_lub<STRESS>(P, A, B, rA, rB, vA, vB, wA, wB, rB, rA);
            if (o.act) {
                if (p == n1) { Fl = add(Fl, o.F); Ml = add(Ml, o.T1); }
                else { Fl = sub(Fl, o.F); Ml = add(Ml, o.T2); }
                if constexpr (STRESS) {
                    for (int e = 0; e < 6; ++e) Sl[e] += p == n1 ? o.S1[e] : o.S2[e];
                    if (o.normal) lubc += 1.0;
                }
            }
        }
        // particle-particle contact (DEM.cpp:1487-1791)
        const double rad = 0.5 * (rA + rB);
        const double x12 = A.x - B.x, y12 = A.y - B.y, z12 = A.z - B.z;
        const double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
        const double urx = vA.x - vB.x, ury = vA.y - vB.y, urz = vA.z - vB.z;
        const double nx = x12 / r12, ny = y12 / r12, nz = z12 / r12;
        const double gn = urx * nx + ury * ny + urz * nz;
        const double del = 2 * rad - r12;
        double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0;
        V fel{0, 0, 0}, fvis{0, 0, 0};  // DEM.cpp:1555-1679
        if (del > 0.0) {
            const double k_ = P.linearStiff, mu = P.frictionCoefPart;
            fnx = k_ * del * nx - dampCoeff * gn * nx;
            fny = k_ * del * ny - dampCoeff * gn * ny;
            fnz = k_ * del * nz - dampCoeff * gn * nz;
            if constexpr (STRESS) {
                fel = V{k_ * del * nx, k_ * del * ny, k_ * del * nz};
                fvis = V{-dampCoeff * gn * nx, -dampCoeff * gn * ny, -dampCoeff * gn * nz};
                // a contact counts where the box-wrapped centres are within 2 rmax (DEM.cpp:1647-1670)
                const double x1 = dbox(A.x, P.demSize[0]), y1 = dbox(A.y, P.demSize[1]), z1 = dbox(A.z, P.demSize[2]);
                const double x2 = dbox(B.x, P.demSize[0]), y2 = dbox(B.y, P.demSize[1]), z2 = dbox(B.z, P.demSize[2]);
                const double zij = z1 - z2, xij = x1 - x2, yij = y1 - y2;
                if (sqrt(zij * zij + yij * yij + xij * xij) <= 2.0 * (P.rmax)) coord += 1.0;
            }
            if (mu != 0.0) {
                const double ox = rA * wA.x + rB * wB.x, oy = rA * wA.y + rB * wB.y, oz = rA * wA.z + rB * wB.z;
                const double vx = urx + ny * oz - nz * oy - gn * nx;
                const double vy = ury + nz * ox - nx * oz - gn * ny;
                const double vz = urz + nx * oy - ny * ox - gn * nz;
                const double xin = xi.x * nx + xi.y * ny + xi.z * nz;
                xi.x -= xin * nx; xi.y -= xin * ny; xi.z -= xin * nz;
                const double kt = 0.2 * k_;
                const double dct = P.kinVisc == 0.0 ? dampCoeff : 0.0;
                ftx = -kt * xi.x - dct * vx; fty = -kt * xi.y - dct * vy; ftz = -kt * xi.z - dct * vz;
                const double fnn = sqrt(fnx * fnx + fny * fny + fnz * fnz);
                const double ftt = sqrt(ftx * ftx + fty * fty + ftz * ftz);
                const double fri = mu * fnn;
                const double tx = ftx / ftt, ty = fty / ftt, tz = ftz / ftt;
                xi.x += vx * P.deltat; xi.y += vy * P.deltat; xi.z += vz * P.deltat;
                if (fri < ftt) {
                    ftx = fri * tx; fty = fri * ty; ftz = fri * tz;
                    xi.x = -(1 / kt) * (ftx + dct * vx);
                    xi.y = -(1 / kt) * (fty + dct * vy);
                    xi.z = -(1 / kt) * (ftz + dct * vz);
                }
                if constexpr (STRESS) fel = add(fel, V{ftx, fty, ftz});
            }
        } else {
            xi = V{0, 0, 0};
        }
        if (p == n1 && slot >= 0) st3(a.xiNew, slot, xi);
        const double dfx = fnx + ftx, dfy = fny + fty, dfz = fnz + ftz;
        const double dtx = fty * nz - ftz * ny, dty = ftz * nx - ftx * nz, dtz = ftx * ny - fty * nx;
        if (p == n1) { Fc = add(Fc, V{dfx, dfy, dfz}); Mc = add(Mc, V{rA * dtx, rA * dty, rA * dtz}); }
        else { Fc = sub(Fc, V{dfx, dfy, dfz}); Mc = add(Mc, V{rB * dtx, rB * dty, rB * dtz}); }
        if constexpr (STRESS) {  // the same tensors for n1 and n2; FElas with the side's sign
            add_arm(Sc, x12, y12, z12, dfx, dfy, dfz);
            add_arm(Se, x12, y12, z12, fel.x, fel.y, fel.z);
            add_arm(Sv, x12, y12, z12, fvis.x, fvis.y, fvis.z);
            Fe = p == n1 ? add(Fe, fel) : sub(Fe, fel);
        }
        // electrostatic repulsion over every listed pair (DEM.cpp:1404-1484)
        {
            double fx, fy, fz;
            if (del <= 0.0) {
                const double e = P.electroForce * exp(del / P.debyeLength);
                fx = e * nx; fy = e * ny; fz = e * nz;
            } else {
                fx = P.electroForce * nx; fy = P.electroForce * ny; fz = P.electroForce * nz;
            }
            if (p == n1) Fr = add(Fr, V{fx, fy, fz}); else Fr = sub(Fr, V{fx, fy, fz});
            if constexpr (STRESS) add_arm(Sr, x12, y12, z12, fx, fy, fz);
        }
    }
    Fc = group_sum(Fc); Mc = group_sum(Mc); Fl = group_sum(Fl); Ml = group_sum(Ml); Fr = group_sum(Fr);
    double Sw[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // contact stress of walls and impactors
    if constexpr (STRESS) {
        for (int e = 0; e < 6; ++e) Sl[e] = group_sum(Sl[e]);
        for (int e = 0; e < 9; ++e) {
            Sc[e] = group_sum(Sc[e]); Sr[e] = group_sum(Sr[e]); Se[e] = group_sum(Se[e]); Sv[e] = group_sum(Sv[e]);
        }
        Fe = group_sum(Fe);
        coord = group_sum(coord);
        lubc = group_sum(lubc);
    }
    if (!valid || lane != 0) return;
    V xiw = ld3(a.xiw, p), xio = ld3(a.xio, p);
    double* oc = a.objContrib + int64_t(p) * 12 * P.nObj;
    double* wc = a.wallContrib + int64_t(p) * 6 * P.nWall;
    for (int c = 0; c < 12 * P.nObj; ++c) oc[c] = 0.0;
    for (int c = 0; c < 6 * P.nWall; ++c) wc[c] = 0.0;
    if (waited) {
        // wall lubrication (DEM.cpp:2525-2672)
        for (int w = 0; w < P.nWall; ++w) {
            const double rad = rp;
            double x12 = (xp.x - P.wallP[w][0]) * P.wallN[w][0] * P.wallN[w][0];
            double y12 = (xp.y - P.wallP[w][1]) * P.wallN[w][1] * P.wallN[w][1];
            double z12 = (xp.z - P.wallP[w][2]) * P.wallN[w][2] * P.wallN[w][2];
            const double r12 = sqrt(x12 * x12 + y12 * y12 + z12 * z12);
            const double dr = r12 - rad;
            const double cut = P.lubCutoff, lub_N = P.lubNormal, lub_T = P.lubTangential, lub_R = P.lubRotation;
            const double h = dmaxd(dr, cut);
            if (!(dr < dmaxd(cut, dmaxd(lub_N, dmaxd(lub_T, lub_R))))) continue;
            const double alpha = M_PI * P.kinVisc * rad * P.deltat;
            double Xa, Ya, Yb1, Yc11, h_c, beta, Xg1 = 0, Yg1 = 0, Yh11 = 0;
            if (dr < dmaxd(cut, lub_N)) {
                Xa = 1.0; h_c = dmaxd(dr, lub_N); Xa *= rad / h - rad / h_c; Xa *= 6.0 * alpha;
                if constexpr (STRESS) { Xg1 = 3.0 / 2.0; Xg1 *= rad / h - rad / h_c; Xg1 *= 4.0 * alpha * rad; }
            } else Xa = 0.0;
            if (dr < dmaxd(cut, lub_T)) {
                Ya = 8.0 / 15.0; Yb1 = -1.0 / 5.0; h_c = dmaxd(dr, lub_T); beta = alpha * log(h_c / h);
                Ya *= 6.0 * beta; Yb1 *= 4.0 * beta * rad;
                if constexpr (STRESS) { Yg1 = 7.0 / 10.0; Yg1 *= 4.0 * beta * rad; }
            } else Ya = Yb1 = 0.0;
            if (dr < dmaxd(cut, lub_R)) {
                Yc11 = 2.0 / 5.0; h_c = dmaxd(dr, lub_R); beta = alpha * log(h_c / h); Yc11 *= 8.0 * beta * rad * rad;
                if constexpr (STRESS) { Yh11 = -1.0 / 10.0; Yh11 *= 8.0 * beta * rad * rad; }
            } else Yc11 = 0.0;
            x12 /= r12; y12 /= r12; z12 /= r12;
            const double umx = vp.x, umy = vp.y, umz = vp.z, w1x = wp.x, w1y = wp.y, w1z = wp.z;
            const double um_r = umx * x12 + umy * y12 + umz * z12;
            const double w1_r = w1x * x12 + w1y * y12 + w1z * z12;
            const double umxrx = umy * z12 - umz * y12, umxry = umz * x12 - umx * z12, umxrz = umx * y12 - umy * x12;
            const double w1xrx = w1y * z12 - w1z * y12, w1xry = w1z * x12 - w1x * z12, w1xrz = w1x * y12 - w1y * x12;
            double Fx = -Xa * x12 * um_r - Ya * (umx - um_r * x12);
            double Fy = -Xa * y12 * um_r - Ya * (umy - um_r * y12);
            double Fz = -Xa * z12 * um_r - Ya * (umz - um_r * z12);
            Fx += -Yb1 * w1xrx; Fy += -Yb1 * w1xry; Fz += -Yb1 * w1xrz;
            double T1x = Yb1 * umxrx, T1y = Yb1 * umxry, T1z = Yb1 * umxrz;
            T1x += -Yc11 * (w1x - w1_r * x12); T1y += -Yc11 * (w1y - w1_r * y12); T1z += -Yc11 * (w1z - w1_r * z12);
            Fl = add(Fl, V{Fx, Fy, Fz}); Ml = add(Ml, V{T1x, T1y, T1z});
            wc[6 * w + 3] -= Fx; wc[6 * w + 4] -= Fy; wc[6 * w + 5] -= Fz;
            if constexpr (STRESS) {  // DEM.cpp:2623-2634 (the wall's own bracketing of the Xg1 term)
                const double xs[3] = {x12, y12, z12}, us[3] = {umx, umy, umz}, w1s[3] = {w1xrx, w1xry, w1xrz};
                const int ci[6] = {0, 1, 2, 1, 2, 0}, cj[6] = {0, 1, 2, 2, 0, 1};
                for (int e = 0; e < 6; ++e) {
                    const int i = ci[e], j = cj[e];
                    double s1 = Xg1 * (um_r * xs[i] * xs[j]) + Yg1 * (xs[i] * us[j] + us[i] * xs[j]);
                    s1 += -Yh11 * (xs[i] * w1s[j] + w1s[i] * xs[j]);
                    Sl[e] += s1;
                }
            }
        }
        // object-particle contact (DEM.cpp:912-1138); history xio shared by all objects
        for (int o = 0; o < P.nObj; ++o) {
            const V ox = ld3(a.ox0, o), ov = ld3(a.ox1, o), ow = ld3(a.ow1, o);
            const double orr = a.or_[o];
            const double radw = 0.5 * (rp + orr);
            const double x1w = xp.x - ox.x, y1w = xp.y - ox.y, z1w = xp.z - ox.z;
            const double r1w = sqrt(x1w * x1w + y1w * y1w + z1w * z1w);
            const double n1wx = x1w / r1w, n1wy = y1w / r1w, n1wz = z1w / r1w;
            const double urx = vp.x - ov.x, ury = vp.y - ov.y, urz = vp.z - ov.z;
            const double delw = (2 * radw) - r1w;
            double fnx = 0, fny = 0, fnz = 0, ftx = 0, fty = 0, ftz = 0;
            double felx = 0, fely = 0, felz = 0, fvisx = 0, fvisy = 0, fvisz = 0;  // DEM.cpp:970-1056
            if (delw > 0.0) {
                const double gn = urx * n1wx + ury * n1wy + urz * n1wz;
                const double k = P.linearStiff, mu = P.frictionCoefPart;
                fnx = k * delw * n1wx - dampCoeff * gn * n1wx;
                fny = k * delw * n1wy - dampCoeff * gn * n1wy;
                fnz = k * delw * n1wz - dampCoeff * gn * n1wz;
                felx = k * delw * n1wx; fely = k * delw * n1wy; felz = k * delw * n1wz;
                fvisx = -dampCoeff * gn * n1wx; fvisy = -dampCoeff * gn * n1wy; fvisz = -dampCoeff * gn * n1wz;
                if (mu != 0.0) {
                    const double o1x = rp * wp.x + orr * ow.x, o1y = rp * wp.y + orr * ow.y, o1z = rp * wp.z + orr * ow.z;
                    const double vx = urx + n1wy * o1z - n1wz * o1y - gn * n1wx;
                    const double vy = ury + n1wz * o1x - n1wx * o1z - gn * n1wy;
                    const double vz = urz + n1wx * o1y - n1wy * o1x - gn * n1wz;
                    const double xin = xio.x * n1wx + xio.y * n1wy + xio.z * n1wz;
                    xio.x -= xin * n1wx; xio.y -= xin * n1wy; xio.z -= xin * n1wz;
                    const double kt = 0.2 * k;
                    // dampCoeffTanw: declared without an initialiser and never assigned
                    // (DEM.cpp:916); taken as zero
                    const double dct = 0.0;
                    ftx = -kt * xio.x; fty = -kt * xio.y; ftz = -kt * xio.z;
                    const double fnn = sqrt(fnx * fnx + fny * fny + fnz * fnz);
                    const double ftt = sqrt(ftx * ftx + fty * fty + ftz * ftz);
                    const double fri = mu * fnn;
                    const double tx = ftx / ftt, ty = fty / ftt, tz = ftz / ftt;
                    if (fri < ftt) {
                        ftx = fri * tx; fty = fri * ty; ftz = fri * tz;
                        // the capped split as written (DEM.cpp:1000-1007): the z term of fel goes
                        // to y, and all three terms of fvis go to x
                        const double fne = sqrt(felx * felx + fely * fely + felz * felz);
                        const double fnv = sqrt(fvisx * fvisx + fvisy * fvisy + fvisz * fvisz);
                        felx += mu * fne * tx; fely += mu * fne * ty; fely += mu * fne * tz;
                        fvisx += mu * fnv * tx; fvisx += mu * fnv * ty; fvisx += mu * fnv * tz;
                        xio.x = -(1 / kt) * (ftx + dct * vx);
                        xio.y = -(1 / kt) * (fty + dct * vy);
                        xio.z = -(1 / kt) * (ftz + dct * vz);
                    } else {
                        felx += -kt * xio.x; fely += -kt * xio.y; felz += -kt * xio.z;
                        fvisx += -dct * vx; fvisy += -dct * vy; fvisz += -dct * vz;
                        xio.x += vx * P.deltat; xio.y += vy * P.deltat; xio.z += vz * P.deltat;
                    }
                }
            } else {
                xio = V{0, 0, 0};
            }
            const V df{fnx + ftx, fny + fty, fnz + ftz};
            const double dtx = fty * n1wz - ftz * n1wy, dty = ftz * n1wx - ftx * n1wz, dtz = ftx * n1wy - fty * n1wx;
            Fc = add(Fc, df);
            Mc = add(Mc, V{rp * dtx, rp * dty, rp * dtz});
            if constexpr (STRESS) {  // DEM.cpp:1091-1135
                add_arm(Sw, x1w, y1w, z1w, df.x, df.y, df.z);
                add_arm(Se, x1w, y1w, z1w, felx, fely, felz);
                add_arm(Sv, x1w, y1w, z1w, fvisx, fvisy, fvisz);
                Fe = add(Fe, V{felx, fely, felz});
            }
            double* c = oc + 12 * o;
            c[0] = -df.x; c[1] = -df.y; c[2] = -df.z;
            c[3] = orr * dtx; c[4] = orr * dty; c[5] = orr * dtz;
        }
        // object-particle lubrication (DEM.cpp:1140-1340)
        for (int o = 0; o < P.nObj; ++o) {
            const LubOut lo = pair_lub<STRESS>(P, xp, ld3(a.ox0, o), rp, a.or_[o], vp, ld3(a.ox1, o), wp, ld3(a.ow1, o), a.or_[o], rp);
            if (!lo.act) continue;
            Fl = add(Fl, lo.F); Ml = add(Ml, lo.T1);
            if constexpr (STRESS)
                for (int e = 0; e < 6; ++e) Sl[e] += lo.S1[e];
            double* c = oc + 12 * o;
            c[6] = -lo.F.x; c[7] = -lo.F.y; c[8] = -lo.F.z;
            c[9] = lo.T2.x; c[10] = lo.T2.y; c[11] = lo.T2.z;
        }
    } else {
        // fictional lid at demSize[1] - 0.3 (DEM.cpp:2157-2296)
        V F, M;
        wall_like_contact(P, dampCoeff, rp, vp, wp, xiw, 0.0, xp.y - (P.demSize[1] - 0.3), 0.0, F, M);
        Fc = add(Fc, F); Mc = add(Mc, M);
    }
    // wall contacts (DEM.cpp:1946-2153), shared history xiw
    for (int w = 0; w < P.nWall; ++w) {
        const double x1w = (xp.x - P.wallP[w][0]) * P.wallN[w][0] * P.wallN[w][0];
        const double y1w = (xp.y - P.wallP[w][1]) * P.wallN[w][1] * P.wallN[w][1];
        const double z1w = (xp.z - P.wallP[w][2]) * P.wallN[w][2] * P.wallN[w][2];
        V F, M, fel, fvis;
        wall_like_contact(P, dampCoeff, rp, vp, wp, xiw, x1w, y1w, z1w, F, M, &fel, &fvis);
        Fc = add(Fc, F); Mc = add(Mc, M);
        wc[6 * w + 0] -= F.x; wc[6 * w + 1] -= F.y; wc[6 * w + 2] -= F.z;
        if constexpr (STRESS) {  // DEM.cpp:2103-2148
            add_arm(Sw, x1w, y1w, z1w, F.x, F.y, F.z);
            add_arm(Se, x1w, y1w, z1w, fel.x, fel.y, fel.z);
            add_arm(Sv, x1w, y1w, z1w, fvis.x, fvis.y, fvis.z);
            Fe = add(Fe, fel);
        }
    }
    if constexpr (STRESS) {
        double* rec = a.stressPart + int64_t(p) * SP_REC;
        for (int e = 0; e < 9; ++e) rec[SP_HYDRO + e] = a.shydro[9 * p + e];
        // SLub in full (m01 = m10 = xy, m02 = m20 = zx, m12 = m21 = yz)
        const double full[9] = {Sl[0], Sl[5], Sl[4], Sl[5], Sl[1], Sl[3], Sl[4], Sl[3], Sl[2]};
        for (int e = 0; e < 9; ++e) {
            rec[SP_LUB + e] = full[e];
            rec[SP_CPAIR + e] = Sc[e];
            rec[SP_CWALL + e] = Sw[e];
            rec[SP_REP + e] = Sr[e];
            rec[SP_ELAS + e] = Se[e];
            rec[SP_VISC + e] = Sv[e];
        }
        rec[SP_FELAS] = Fe.x; rec[SP_FELAS + 1] = Fe.y; rec[SP_FELAS + 2] = Fe.z;
        rec[SP_COORD] = coord; rec[SP_LUBC] = lubc;
    }
    st3(a.FContact, p, Fc); st3(a.MContact, p, Mc); st3(a.FLub, p, Fl); st3(a.MLub, p, Ml); st3(a.FRep, p, Fr);
    st3(a.xiw, p, xiw); st3(a.xio, p, xio);
}

// Deterministic column sums of the per-particle object and wall contributions (row-major,
// ncol columns), in two levels: block b sums rows [b*rows, (b+1)*rows) of every column
// (threads stride over the flat range, so reads are coalesced) into part[b][c]; one block then
// sums the block partials of each column in block order. Blocks [0, RED_BLOCKS) take the
// objects, [RED_BLOCKS, 2 RED_BLOCKS) the walls.
constexpr int RED_BLOCKS = 64, RED_BS = 256, RED_MAXCOL = 12 * MAX_OBJECTS;
__device__ inline void k_dem_reduce_a(const double* oc, int nco, const double* wcn, int ncw, int nPart,
                                      double* part) {
    __shared__ double red[RED_BS];
    const int role = int(blockIdx.x) / RED_BLOCKS, bx = int(blockIdx.x) % RED_BLOCKS;
    const double* contrib = role == 0 ? oc : wcn;
    const int ncol = role == 0 ? nco : ncw;
    if (ncol == 0) return;
    const int rows = (nPart + RED_BLOCKS - 1) / RED_BLOCKS;
    const int64_t beg = int64_t(bx) * rows * ncol;
    const int64_t end = min(int64_t(nPart), int64_t(bx + 1) * rows) * ncol;
    // thread t handles flat indices beg + t + k*step with step a multiple of ncol, so its column
    // is fixed: c = (t + beg) % ncol
    const int step = (RED_BS / ncol) * ncol;
    double acc = 0.0;
    const int t = threadIdx.x;
    if (t < step) {
        // eight loads in flight, added in index order (the same sum as one at a time)
        int64_t i = beg + t;
        for (; i + 7 * int64_t(step) < end; i += 8 * int64_t(step)) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = contrib[i + u * int64_t(step)];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; i < end; i += step) acc += contrib[i];
    }
    red[t] = t < step ? acc : 0.0;
    __syncthreads();
    if (t < ncol) {
        double s = 0.0;
        for (int u = t; u < step; u += ncol) s += red[u];
        part[(int64_t(role) * RED_BLOCKS + bx) * RED_MAXCOL + t] = s;
    }
}
// One block (the last block of the integration launch, which runs beside the particles' blocks:
// neither reads what the other writes): the second level of the sums (the block partials of each
// column in block order; or zeros when no particle contributes), then, on thread 0, objects:
// totals and integration (DEM.cpp:786-790, 2742-2758) and walls: totals (DEM.cpp:773-775).
constexpr int OBJ_THREADS = 128;
static_assert(RED_BLOCKS % 16 == 0, "column_sum loads sixteen partials at a time");
// the RED_BLOCKS block partials of one column (stride RED_MAXCOL), added in block order; the loads
// are issued sixteen at a time
__device__ inline double column_sum(const double* col) {
    double s = 0.0;
#pragma unroll
    for (int b0 = 0; b0 < RED_BLOCKS; b0 += 16) {
        double v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = col[int64_t(b0 + u) * RED_MAXCOL];
#pragma unroll
        for (int u = 0; u < 16; ++u) s += v[u];
    }
    return s;
}
static_assert(OBJ_THREADS >= 12 * MAX_OBJECTS && OBJ_THREADS >= 6 * MAX_WALLS, "one thread per column");
__device__ inline void dem_objects(DemParams P, DemState d, const double* part, int sumsFromParts, double* objSum,
                                   double* wallSum, int demTimeStep) {
#pragma clang fp contract(off)
    __shared__ double sObj[12 * MAX_OBJECTS], sWall[6 * MAX_WALLS];
    const int t = threadIdx.x;
    const int nco = 12 * P.nObj, ncw = 6 * P.nWall;
    if (t < 12 * MAX_OBJECTS) {
        double s = 0.0;
        if (sumsFromParts && t < nco) s = column_sum(part + t);
        sObj[t] = s;
        objSum[t] = s;
    }
    if (t < 6 * MAX_WALLS) {
        double s = 0.0;
        if (sumsFromParts && t < ncw) s = column_sum(part + int64_t(RED_BLOCKS) * RED_MAXCOL + t);
        sWall[t] = s;
        wallSum[t] = s;
    }
    __syncthreads();
    if (t != 0) return;
    objSum = sObj;
    wallSum = sWall;
    // body-spring forces (DEM.cpp:798-869), on the positions before this step's integration
    double spring[MAX_OBJECTS];  // y component of each object's FSpring
    for (int o = 0; o < MAX_OBJECTS; ++o) spring[o] = 0.0;
    const bool waited = demTimeStep >= P.objWaitTime;
    for (int b = 0; b < P.nBody; ++b) {
        double* B = d.body + b * LBDEM_B_REC;
        double fsB = 0.0;
        if (waited) {
            for (int o = 0; o < P.nObj; ++o) {
                const double L = B[LBDEM_B_X0 + 1] - d.ox0[3 * o + 1];
                B[LBDEM_B_L] = L;
                const double rx = B[LBDEM_B_X0] - d.ox0[3 * o], ry = L, rz = B[LBDEM_B_X0 + 2] - d.ox0[3 * o + 2];
                const double rn = sqrt(rx * rx + ry * ry + rz * rz);
                const double ny = ry / rn;
                const double k = B[LBDEM_B_K], del = B[LBDEM_B_L0] - L, sigma = B[LBDEM_B_R];
                const double bodyDamp = 2.0 * sqrt(B[LBDEM_B_M] * k) * P.springDamp;
                const double objDamp = 2.0 * sqrt(d.om[o] * k) * P.springDamp;
                double fs, fdo = -d.ox1[3 * o + 1] * objDamp, fdb = -B[LBDEM_B_X1 + 1] * bodyDamp;
                if (P.nonLinType == 0) {
                    fs = k * del * ny;
                    fdo = 0.0;
                    fdb = 0.0;
                } else if (P.nonLinType == 1) {
                    fs = (k * del) * ny;
                } else if (P.nonLinType == 2) {
                    fs = (k * del + P.k13 * k * (del * del * del)) * ny;
                } else if (P.nonLinType == 3) {
                    fs = (k * del + (24.0 * P.epsilon / rn) * 2.0 * (pow((sigma / rn), 12.0))) * ny;
                } else {  // 4: WCA repulsion inside 2^(1/6) sigma
                    if (rn <= pow(2.0, 1.0 / 6.0) * sigma)
                        fs = (k * del + (24.0 * P.epsilon / rn) * ((2.0 * pow(sigma / rn, 12.0)) - (pow(sigma / rn, 6.0)))) * ny;
                    else
                        fs = (k * del) * ny;
                }
                spring[o] -= fs;
                fsB += fs;
                spring[o] += fdo;
                fsB += fdb;
            }
        }
        // totals (DEM.cpp:792-794): FTotal = FGrav + FSpring; integration (DEM.cpp:2730-2740)
        const double FT[3] = {B[LBDEM_B_FGRAV] + 0.0, B[LBDEM_B_FGRAV + 1] + fsB, B[LBDEM_B_FGRAV + 2] + 0.0};
        B[LBDEM_B_FSPRING] = 0.0; B[LBDEM_B_FSPRING + 1] = fsB; B[LBDEM_B_FSPRING + 2] = 0.0;
        for (int c = 0; c < 3; ++c) B[LBDEM_B_FTOTAL + c] = FT[c];
        if (waited) {
            const double dt = P.deltat, im = 1.0 / B[LBDEM_B_M];
            for (int c = 0; c < 3; ++c) B[LBDEM_B_X0 + c] += dt * B[LBDEM_B_X1 + c] + dt * dt * FT[c] * im;
            for (int c = 0; c < 3; ++c) B[LBDEM_B_X1 + c] += dt * FT[c] * im;
        }
    }
    for (int o = 0; o < P.nObj; ++o) {
        const double* s = objSum + 12 * o;
        const V Fc{s[0], s[1], s[2]}, Mc{s[3], s[4], s[5]}, Fl{s[6], s[7], s[8]}, Ml{s[9], s[10], s[11]};
        st3(d.oFContact, o, Fc); st3(d.oMContact, o, Mc); st3(d.oFLub, o, Fl); st3(d.oMLub, o, Ml);
        const V FH = ld3(d.oFHydro, o), MH = ld3(d.oMHydro, o), FG = ld3(d.oFGrav, o);
        const V FT = add(add(add(add(FH, Fc), Fl), FG), V{0, spring[o], 0});
        const V MT = add(add(MH, Mc), Ml);
        st3(d.oFTotal, o, FT); st3(d.oMTotal, o, MT);
        if (demTimeStep >= P.objWaitTime) {
            const double dt = P.deltat, im = d.oim[o], iI = d.oiI[o];
            double* x0 = d.ox0 + 3 * o; double* x1 = d.ox1 + 3 * o; double* w1 = d.ow1 + 3 * o;
            const double F[3] = {FT.x, FT.y, FT.z}, M[3] = {MT.x, MT.y, MT.z};
            for (int c = 0; c < 3; ++c) x0[c] += dt * x1[c] + dt * dt * F[c] * im;
            for (int c = 0; c < 3; ++c) x1[c] += dt * F[c] * im;
            for (int c = 0; c < 3; ++c) w1[c] += dt * M[c] * iI;
        }
        ObjLB ol;
        for (int c = 0; c < 3; ++c) { ol.c[c] = d.ox0[3 * o + c] / P.unitLength; ol.u[c] = d.ox1[3 * o + c] / P.unitSpeed; }
        ol.r = d.or_[o] / P.unitLength;
        ol.pad = 0.0;
        d.olb[o] = ol;
    }
    for (int w = 0; w < P.nWall; ++w) {
        const double* s = wallSum + 6 * w;
        const V Fc{s[0], s[1], s[2]}, Fl{s[3], s[4], s[5]};
        st3(d.wFContact, w, Fc); st3(d.wFLub, w, Fl);
        st3(d.wFTotal, w, add(add(add(ld3(d.wFHydro, w), Fc), Fl), V{0, 0, 0}));
    }
}

// particle p: totals, gravity switch-off, integration, periodic wrap, lattice view; returns |v|^2
__device__ inline double integrate_particle(const DemParams& P, const DemState& d, int demTimeStep, int p) {
#pragma clang fp contract(off)
    const V FT = add(add(add(add(ld3(d.FHydro, p), ld3(d.FContact, p)), ld3(d.FLub, p)), ld3(d.FGrav, p)), ld3(d.FRep, p));
    const V MT = add(add(ld3(d.MHydro, p), ld3(d.MContact, p)), ld3(d.MLub, p));
    st3(d.FTotal, p, FT); st3(d.MTotal, p, MT);
    if (demTimeStep <= P.objWaitTime) st3(d.FGrav, p, V{0, 0, 0});
    const double dt = P.deltat, im = d.im[p], iI = d.iI[p];
    double* x0 = d.x0 + 3 * p; double* x1 = d.x1 + 3 * p; double* w1 = d.w1 + 3 * p;
    const double F[3] = {FT.x, FT.y, FT.z}, M[3] = {MT.x, MT.y, MT.z};
    for (int c = 0; c < 3; ++c) x0[c] += dt * x1[c] + dt * dt * F[c] * im;
    for (int c = 0; c < 3; ++c) x1[c] += dt * F[c] * im;
    for (int c = 0; c < 3; ++c) w1[c] += dt * M[c] * iI;
    if (P.stress) {  // DEM.cpp:2705, 2721-2725 (before the periodic wrap, which follows the integration)
        for (int c = 0; c < 3; ++c) d.disp[3 * p + c] = d.xi0[3 * p + c] - x0[c];
        if (demTimeStep == P.objWaitTime)
            for (int c = 0; c < 3; ++c) d.xi0[3 * p + c] = x0[c];
    }
    for (int c = 0; c < 3; ++c)
        if (P.periodic[c]) {
            if (x0[c] > P.demSize[c]) x0[c] -= P.demSize[c];
            if (x0[c] < 0.0) x0[c] += P.demSize[c];
        }
    PartLB v;
    for (int c = 0; c < 3; ++c) {
        v.c[c] = x0[c] / P.unitLength;
        v.v[c] = x1[c] / P.unitSpeed;
        v.w[c] = w1[c] / P.unitAngVel;
    }
    v.r = d.r[p] / P.unitLength;
    d.plb[p] = v;
    return x1[0] * x1[0] + x1[1] * x1[1] + x1[2] * x1[2];
}
// Blocks [0, nb) integrate the particles (OBJ_THREADS per block), block nb does the objects and
// walls (dem_objects); later blocks (a batch launch sized for a larger box) return.
__device__ inline void k_dem_integrate(DemParams P, DemState d, const double* part, int sumsFromParts, int demTimeStep,
                                       int nb) {
#pragma clang fp contract(off)
    if (int(blockIdx.x) >= nb) {
        if (int(blockIdx.x) == nb) dem_objects(P, d, part, sumsFromParts, d.objSum, d.wallSum, demTimeStep);
        return;
    }
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double v2 = 0.0;  // |v|^2 after the update, for the next step's neighbour-list trigger
    if (p < P.nPart) v2 = integrate_particle(P, d, demTimeStep, p);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v2, off, 64);
        v2 = o > v2 ? o : v2;
    }
    if (__lane_id() == 0 && v2 > 0.0)
        atomicMax(d.blockMax + blockIdx.x % VEL_SLOTS, (unsigned long long)__double_as_longlong(v2));
}

__device__ inline void k_dem_views(DemParams P, DemState d) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < P.nPart) {
        PartLB v;
        for (int c = 0; c < 3; ++c) {
            v.c[c] = d.x0[3 * p + c] / P.unitLength;
            v.v[c] = d.x1[3 * p + c] / P.unitSpeed;
            v.w[c] = d.w1[3 * p + c] / P.unitAngVel;
        }
        v.r = d.r[p] / P.unitLength;
        d.plb[p] = v;
    }
    if (p < P.nObj) {
        ObjLB ol;
        for (int c = 0; c < 3; ++c) { ol.c[c] = d.ox0[3 * p + c] / P.unitLength; ol.u[c] = d.ox1[3 * p + c] / P.unitSpeed; }
        ol.r = d.or_[p] / P.unitLength;
        ol.pad = 0.0;
        d.olb[p] = ol;
    }
}


void dem_refresh_views(const DemParams& p, DemState& d, hipStream_t s) {
    const int n = p.nPart > p.nObj ? p.nPart : p.nObj;
    if (n > 0) launch<&k_dem_views, 1024, 1>((n + 255) / 256, 256, s, p, d);
}

static NbrArgs nbr_args(const DemParams& p, DemState& d, DevScalars* sc) {
    NbrArgs a;
    a.p = p; a.x0 = d.x0; a.r = d.r;
    a.cellCount = d.cellCount; a.cellStart = d.cellStart; a.cellFill = d.cellFill; a.cellItems = d.cellItems;
    a.cellKey = d.cellKey; a.nbrCount = d.nbrCount; a.nbrList = d.nbrList; a.nbrSlot = d.nbrSlot;
    a.sc = sc;
    return a;
}

void dem_neighbor_list(const DemParams& p, DemState& d, DevScalars* sc, bool force, hipStream_t s) {
    launch<&k_maxdisp, 1024, 1>(1, VEL_SLOTS, s, d.blockMax, p.deltat, sc, force ? 1 : 0);
    if (p.nPart <= 0) return;
    NbrArgs a = nbr_args(p, d, sc);
    const int pb = (p.nPart + 127) / 128;
    if (p.nlx > 0 && p.nlz > 0) {
        const int ncell = p.nlx * p.nly * p.nlz;
        launch<&k_cells_assign, 1024, 1>(pb, 128, s, a);
        launch<&k_cells_scan, SCAN_THREADS, 1>(1, SCAN_THREADS, s, a, ncell);
        launch<&k_cells_fill, 1024, 1>(pb, 128, s, a);
    }
    launch<&k_nbr_build, NBR_BLOCK, 1>((p.nPart + NBR_GROUPS - 1) / NBR_GROUPS, NBR_BLOCK, s, a);
    const int eb = int((int64_t(p.nPart) * SLOT_LANES + 255) / 256);  // SLOT_LANES threads per particle
    launch<&k_nbr_slots, 256, 1>(eb, 256, s, a, d.xiCur, d.nbrCountOld, d.nbrListOld);
}

bool dem_step(const DemParams& p, DemState& d, DevScalars* sc, int demTimeStep, hipStream_t s, hipEvent_t hydroReady) {
    if (p.nPart > 0) {
        ForceArgs f;
        f.p = p; f.x0 = d.x0; f.x1 = d.x1; f.w1 = d.w1; f.r = d.r;
        f.FContact = d.FContact; f.MContact = d.MContact; f.FLub = d.FLub; f.MLub = d.MLub; f.FRep = d.FRep;
        f.xiw = d.xiw; f.xio = d.xio; f.nbrCount = d.nbrCount; f.nbrList = d.nbrList; f.nbrSlot = d.nbrSlot;
        f.xiOld = d.xiCur; f.xiNew = d.xiNext;
        f.ox0 = d.ox0; f.ox1 = d.ox1; f.ow1 = d.ow1; f.or_ = d.or_;
        f.objContrib = d.objContrib; f.wallContrib = d.wallContrib; f.demTimeStep = demTimeStep;
        f.stressPart = d.stressPart; f.shydro = d.shydro;
        if (p.stress) launch<&k_dem_forces<true>, 256, 1>((p.nPart * DEM_LANES + 255) / 256, 256, s, f, p.dampCoeff);
        else launch<&k_dem_forces<false>, 256, 1>((p.nPart * DEM_LANES + 255) / 256, 256, s, f, p.dampCoeff);
        if (p.nObj > 0 || p.nWall > 0)
            launch<&k_dem_reduce_a, RED_BS, 1>(2 * RED_BLOCKS, RED_BS, s, d.objContrib, 12 * p.nObj, d.wallContrib, 6 * p.nWall, p.nPart, d.redPart);
        double* t = d.xiCur; d.xiCur = d.xiNext; d.xiNext = t;
    }
    const int fromParts = (p.nPart > 0 && (p.nObj > 0 || p.nWall > 0)) ? 1 : 0;  // else zero sums
    const int nb = (p.nPart + OBJ_THREADS - 1) / OBJ_THREADS;
    // the integration reads the hydrodynamic sums (particles, impactors, walls)
    const bool ok = !hydroReady || hipStreamWaitEvent(s, hydroReady, 0) == hipSuccess;
    launch<&k_dem_integrate, OBJ_THREADS, 1>(nb + 1, OBJ_THREADS, s, p, d, (const double*)d.redPart, fromParts,
                                             demTimeStep, nb);
    return ok;
}

}  // namespace lbdem
