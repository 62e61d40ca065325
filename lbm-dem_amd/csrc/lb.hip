// Lattice kernels of the coupled step on gfx950:
//   * k_fused      — reconstruct + particle interior forcing + BGK/Guo collision + AA streaming
//                    with gas / wall / moving-object boundary conditions resolved by the storing
//                    cell (reference LB.cpp:1261-1544, 1901-1978, node.cpp:75-207)
//   * free-surface mass exchange and interface conversion (LB.cpp:1548-1852)
//   * particle flags and the moving-object (impactor) node update (LB.cpp:683-711, 797-943,
//     1875-1899, 1980-2029; findNewSolid, LB.cpp:2032-2100, changes no flag)
#include "lbdem_internal.hpp"

#include <climits>

namespace lbdem {

template <typename S> struct Types;
template <> struct Types<double> { using C = double; using M = double; };
template <> struct Types<float> { using C = float; using M = float; };
template <> struct Types<__half> { using C = float; using M = float; };

__device__ inline bool is_active(uint8_t t) { t &= T_BASE; return t == T_FLUID || t == T_IF; }

// insideSphere with the reference's rounding (myvector.cpp:171-176): no contraction
__device__ inline bool inside_sphere(double px, double py, double pz, double cx, double cy, double cz, double r) {
#pragma clang fp contract(off)
    const double dx = px - cx, dy = py - cy, dz = pz - cz;
    const double s = dx * dx + dy * dy + dz * dz;
    return s < r * r;
}

__device__ inline void cell_xyz(const Geom& g, int64_t i, int& x, int& y, int& z) {
    z = int(i / (int64_t(g.X) * g.Y));
    const int64_t rem = i - int64_t(z) * g.X * g.Y;
    y = int(rem / g.X);
    x = int(rem - int64_t(y) * g.X);
}

// Per-axis index offsets of the wrapped neighbour rule: nb(q) = i + ox[cx+1] + oy[cy+1] + oz[cz+1]
struct Offs {
    int64_t ox[3], oy[3], oz[3];
    __device__ Offs(const Geom& g, int x, int y, int z) {
        const int64_t sy = g.X, sz = int64_t(g.X) * g.Y;
        ox[1] = 0; oy[1] = 0; oz[1] = 0;
        ox[0] = (x - 1 == 0 && g.per[0]) ? int64_t(g.X - 3) : -1;
        ox[2] = (x + 1 == g.X - 1 && g.per[1]) ? -int64_t(g.X - 3) : 1;
        oy[0] = (y - 1 == 0 && g.per[2]) ? int64_t(g.Y - 3) * sy : -sy;
        oy[2] = (y + 1 == g.Y - 1 && g.per[3]) ? -int64_t(g.Y - 3) * sy : sy;
        oz[0] = (z - 1 == 0 && g.per[4]) ? int64_t(g.Z - 3) * sz : -sz;
        oz[2] = (z + 1 == g.Z - 1 && g.per[5]) ? -int64_t(g.Z - 3) * sz : sz;
    }
    __device__ int64_t d(int q) const { return ox[cxq(q) + 1] + oy[cyq(q) + 1] + oz[czq(q) + 1]; }
};

// slot holding the post-streaming population q of cell i in layout L (0 natural, 1 swapped)
__device__ inline int64_t load_slot(const Geom& g, int64_t i, const Offs& o, int q, int L) {
    return L == 0 ? int64_t(q) * g.N + i : int64_t(oppq(q)) * g.N + i + o.d(oppq(q));
}

template <typename C>
__device__ inline void equilibrium_offsets(C n, C ux, C uy, C uz, C* geq) {
    const C usq = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C cu = C(cxq(q)) * ux + C(cyq(q)) * uy + C(czq(q)) * uz;
        geq[q] = C(wq(q)) * ((n - C(1)) + n * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
    }
}

// ------------------------------------------------------------------------------------------
// initial populations (node::initialize, node.cpp:26-73) for fluid and interface nodes
// ------------------------------------------------------------------------------------------
template <typename S>
__global__ void k_init_fields(Geom g, S* f, const uint8_t* type, typename Types<S>::M* mass,
                              typename Types<S>::C* aux, Consts k) {
    using C = typename Types<S>::C;
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= g.N) return;
    const uint8_t t = type[i] & T_BASE;
    if (t != T_FLUID && t != T_IF) return;
    const double n = k.rho0;
    const double ux = k.lbF[0] / 2.0 / n + k.initVel[0];
    const double uy = k.lbF[1] / 2.0 / n + k.initVel[1];
    const double uz = k.lbF[2] / 2.0 / n + k.initVel[2];
    const double usq = ux * ux + uy * uy + uz * uz;
    for (int q = 0; q < Q; ++q) {
        const double vu = ux * cxq(q) + uy * cyq(q) + uz * czq(q);
        const double feq = wq(q) * n * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
        StoreT<S>::st(f + int64_t(q) * g.N + i, C(feq - wq(q)));
    }
    mass[i] = typename Types<S>::M(t == T_FLUID ? k.rho0 : 0.5 * k.rho0);
    if (t == T_IF) {
        aux[i] = C(n);
        aux[g.N + i] = C(ux); aux[2 * g.N + i] = C(uy); aux[3 * g.N + i] = C(uz);
        aux[4 * g.N + i] = C(0); aux[5 * g.N + i] = C(0); aux[6 * g.N + i] = C(0);
    }
}

void lb_init_fields(const Box& b, const Consts& k, hipStream_t s) {
    const int bs = 256;
    const int nb = int((b.g.N + bs - 1) / bs);
    if (b.prec == LBDEM_PREC_FP64)
        hipLaunchKernelGGL(k_init_fields<double>, nb, bs, 0, s, b.g, (double*)b.f, b.type, (double*)b.mass, (double*)b.aux, k);
    else if (b.prec == LBDEM_PREC_FP32)
        hipLaunchKernelGGL(k_init_fields<float>, nb, bs, 0, s, b.g, (float*)b.f, b.type, (float*)b.mass, (float*)b.aux, k);
    else
        hipLaunchKernelGGL(k_init_fields<__half>, nb, bs, 0, s, b.g, (__half*)b.f, b.type, (float*)b.mass, (float*)b.aux, k);
}

// ------------------------------------------------------------------------------------------
// fused collide + stream
// ------------------------------------------------------------------------------------------
constexpr int FX = 64, FY = 4, FZ = 8;  // tile: 64x4 threads, each walks FZ cells along z
constexpr int HS = 64;                  // per-block particle hash slots
constexpr int PART_STRIDE = PART_STRIDE_;  // per-block partials: obj (7*MAX_OBJ_PART) + walls 18
constexpr int MAX_OBJ_PART = 2;

template <typename S>
struct FusedArgs {
    Geom g;
    S* f;
    uint8_t* type;
    const int32_t* solid;
    typename Types<S>::M* mass;
    typename Types<S>::C* aux;
    int32_t* ifList;
    int ifCap;
    const PartLB* plb;
    const ObjLB* olb;
    double* hacc;
    double* blockPart;
    DevScalars* sc;
    int nbx, nby, nbz;
    Consts k;
};

// first lane of the run of equal keys that contains this lane
__device__ inline int run_start(int key, int lane) {
    const int prev = __shfl_up(key, 1, 64);
    int start = (lane == 0 || prev != key) ? lane : 0;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(start, off, 64);
        if (lane >= off && o > start) start = o;
    }
    return start;
}
// segmented inclusive sum over the lanes [start, lane] of the wavefront
template <typename C>
__device__ inline C seg_sum(C v, int start, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const C o = __shfl_up(v, off, 64);
        if (lane - off >= start) v += o;
    }
    return v;
}

// boundary link of the storing cell (LB.cpp:1390-1453): value reflected into the slot the cell
// itself loads next step; momentum exchange into the block partials
template <typename S, bool DIAG>
__device__ inline typename Types<S>::C boundary_value(const FusedArgs<S>& a, int j, int32_t yj, uint8_t ty,
                                                     typename Types<S>::C gj, typename Types<S>::C n,
                                                     typename Types<S>::C sx, typename Types<S>::C sy,
                                                     typename Types<S>::C sz, typename Types<S>::C usq, double mi,
                                                     double* sPart) {
    using C = typename Types<S>::C;
    const Consts& k = a.k;
    if (ty == T_GAS) {
        const C cu = C(cxq(j)) * sx + C(cyq(j)) * sy + C(czq(j)) * sz;
        return -gj + C(wq(j)) * (C(9) * (cu * cu) - C(3) * usq);
    }
    if (ty == T_WALL) {
        if (DIAG && k.nWall > 0) {
            int wi = a.solid[yj];
            wi = (wi >= 0 && wi < k.nWall) ? wi : 0;
            const double v2 = 2.0 * double(gj);
            if (cxq(j)) atomicAdd(&sPart[7 * MAX_OBJ_PART + 3 * wi + 0], v2 * cxq(j));
            if (cyq(j)) atomicAdd(&sPart[7 * MAX_OBJ_PART + 3 * wi + 1], v2 * cyq(j));
            if (czq(j)) atomicAdd(&sPart[7 * MAX_OBJ_PART + 3 * wi + 2], v2 * czq(j));
        }
        return gj;
    }
    if (ty == T_OBJ) {
        const int nObj = k.nObj;
        int ov = 0;
        if (nObj > 1) {
            int yx, yy, yz;
            cell_xyz(a.g, yj, yx, yy, yz);
            for (int oo = 0; oo < nObj; ++oo)
                if (inside_sphere(yx, yy, yz, a.olb[oo].c[0], a.olb[oo].c[1], a.olb[oo].c[2], a.olb[oo].r)) ov = oo;
        }
        const ObjLB& OV = a.olb[ov];
        const double cuo = OV.u[0] * cxq(j) + OV.u[1] * cyq(j) + OV.u[2] * czq(j);
        const C BBi = C(6.0) * n * C(wq(j)) * C(cuo);
        const int oi = a.solid[yj];
        if (oi >= 0 && oi < nObj && oi < MAX_OBJ_PART) {
            int yx, yy, yz;
            cell_xyz(a.g, yj, yx, yy, yz);
            const double fm = 2.0 * double(gj) - double(BBi);
            const double Fb[3] = {fm * cxq(j), fm * cyq(j), fm * czq(j)};
            const double rr[3] = {yx - a.olb[oi].c[0], yy - a.olb[oi].c[1], yz - a.olb[oi].c[2]};
            double* sp = &sPart[7 * oi];
            atomicAdd(sp + 0, Fb[0]); atomicAdd(sp + 1, Fb[1]); atomicAdd(sp + 2, Fb[2]);
            atomicAdd(sp + 3, rr[1] * Fb[2] - rr[2] * Fb[1]);
            atomicAdd(sp + 4, rr[2] * Fb[0] - rr[0] * Fb[2]);
            atomicAdd(sp + 5, rr[0] * Fb[1] - rr[1] * Fb[0]);
        }
        atomicAdd(&sPart[6], double(BBi) * mi);
        return gj - BBi;
    }
    atomicOr(&a.sc->typeError, 1);
    return gj;
}

// plain index offset of direction q (valid for cells without the edge bit)
__device__ inline int32_t plain_off(int q, int32_t sy, int32_t sz) { return cxq(q) + cyq(q) * sy + czq(q) * sz; }

template <typename S, int PAR>
__device__ inline void load_cell(const FusedArgs<S>& a, int32_t i, uint8_t t, int x, int y, int z,
                                 typename Types<S>::C* gq) {
    const int32_t N = int32_t(a.g.N), sy = a.g.X, sz = a.g.X * a.g.Y;
    if (PAR == 0) {
#pragma unroll
        for (int q = 0; q < Q; ++q) gq[q] = StoreT<S>::ld(a.f + q * N + i);
    } else if (!(t & T_EDGE)) {
#pragma unroll
        for (int q = 0; q < Q; ++q) gq[q] = StoreT<S>::ld(a.f + oppq(q) * N + i + plain_off(oppq(q), sy, sz));
    } else {
        const Offs o(a.g, x, y, z);
#pragma unroll
        for (int q = 0; q < Q; ++q) gq[q] = StoreT<S>::ld(a.f + oppq(q) * N + i + int32_t(o.d(oppq(q))));
    }
}

template <typename S, int PAR, bool FSURF, bool DIAG>
__global__ __launch_bounds__(FX* FY) void k_fused(FusedArgs<S> a) {
    using C = typename Types<S>::C;
    using M = typename Types<S>::M;
    __shared__ int hkey[HS];
    __shared__ double hval[HS][7];
    __shared__ double sPart[PART_STRIDE];
    const int tid = threadIdx.x + FX * threadIdx.y;
    for (int s = tid; s < HS; s += FX * FY) {
        hkey[s] = -1;
        for (int c = 0; c < 7; ++c) hval[s][c] = 0.0;
    }
    if (tid < PART_STRIDE) sPart[tid] = 0.0;
    __syncthreads();

    // XCD-aware tile order: consecutive tiles of the box go to the same XCD
    const int nblk = a.nbx * a.nby * a.nbz;
    const int b0 = blockIdx.x;
    const int per = (nblk + 7) / 8;
    const int tile = (b0 % 8) * per + b0 / 8;
    const bool tile_ok = tile < nblk;
    const int bx = tile % a.nbx, by = (tile / a.nbx) % a.nby, bz = tile / (a.nbx * a.nby);
    const int x = bx * FX + threadIdx.x, y = by * FY + threadIdx.y;
    const Geom& g = a.g;
    const int32_t N = int32_t(g.N), SY = g.X, SZ = g.X * g.Y;
    const int lane = __lane_id();
    const Consts& k = a.k;
    const bool col_ok = tile_ok && x < g.X && y < g.Y;
    const int z0 = bz * FZ;

    // prefetch of the first cell of the column
    C gc[Q], gn[Q];
    uint8_t tc = 0;
    int32_t ic = 0;
    bool actc = false;
    if (col_ok && z0 < g.Z) {
        ic = int32_t(cidx(g, x, y, z0));
        tc = a.type[ic];
        actc = is_active(tc);
        if (actc) load_cell<S, PAR>(a, ic, tc, x, y, z0, gc);
    }
    for (int zz = 0; zz < FZ; ++zz) {
        const int z = z0 + zz;
        // issue the loads of the next cell before this cell's stores
        uint8_t tn = 0;
        int32_t in = 0;
        bool actn = false;
        if (zz + 1 < FZ && col_ok && z + 1 < g.Z) {
            in = ic + SZ;
            tn = a.type[in];
            actn = is_active(tn);
            if (actn) load_cell<S, PAR>(a, in, tn, x, y, z + 1, gn);
        }
        const bool act = actc;
        const int32_t i = ic;
        const uint8_t t = tc;
        int hkeyv = -1;
        C dfx = 0, dfy = 0, dfz = 0, dmx = 0, dmy = 0, dmz = 0, dvol = 0;
        if (act) {
            // reconstruct (node.cpp:75-93)
            C sum = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                sum += gc[q];
                jx += C(cxq(q)) * gc[q]; jy += C(cyq(q)) * gc[q]; jz += C(czq(q)) * gc[q];
            }
            const C n = C(1) + sum;
            const C ux = jx / n, uy = jy / n, uz = jz / n;
            // interior forcing of particle nodes (LB.cpp:1919-1962)
            C hx = 0, hy = 0, hz = 0;
            // mass is needed by particle nodes (LB.cpp:1935) and object links (LB.cpp:1453) only
            const double mi = (t & (T_PMASK | T_EDGE)) ? double(a.mass[i]) : 0.0;
            if (t & T_PMASK) {
                const int sp = a.solid[i];
                const PartLB P = a.plb[sp];
                const C rx = C(double(x) - P.c[0]), ry = C(double(y) - P.c[1]), rz = C(double(z) - P.c[2]);
                const C wx = C(P.w[0]), wy = C(P.w[1]), wz = C(P.w[2]);
                const C lvx = C(P.v[0]) + (wy * rz - wz * ry);
                const C lvy = C(P.v[1]) + (wz * rx - wx * rz);
                const C lvz = C(P.v[2]) + (wx * ry - wy * rx);
                const C lf = C(mi) / n;
                dfx = lf * (ux - lvx); dfy = lf * (uy - lvy); dfz = lf * (uz - lvz);
                hx = -dfx; hy = -dfy; hz = -dfz;
                dmx = ry * dfz - rz * dfy; dmy = rz * dfx - rx * dfz; dmz = rx * dfy - ry * dfx;
                dvol = C(mi);
                hkeyv = sp;
            }
            // collision (node.cpp:95-111, 181-207)
            const C Fx = C(k.lbF[0]) + hx, Fy = C(k.lbF[1]) + hy, Fz = C(k.lbF[2]) + hz;
            const C sx = ux + C(0.5) * Fx / n, sy = uy + C(0.5) * Fy / n, sz = uz + C(0.5) * Fz / n;
            const C usq = sx * sx + sy * sy + sz * sz;
            const C uF = sx * Fx + sy * Fy + sz * Fz;
            const C om = C(k.omega), omf = C(k.omegaf);
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const C cu = C(cxq(q)) * sx + C(cyq(q)) * sy + C(czq(q)) * sz;
                const C cF = C(cxq(q)) * Fx + C(cyq(q)) * Fy + C(czq(q)) * Fz;
                const C geq = C(wq(q)) * ((n - C(1)) + n * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
                gc[q] = gc[q] + om * (geq - gc[q]) + omf * C(wq(q)) * (C(9) * cu * cF + C(3) * (cF - uF));
            }
            // streaming (LB.cpp:1375-1528)
            StoreT<S>::st(a.f + i, gc[0]);
            if (!(t & T_EDGE)) {
#pragma unroll
                for (int j = 1; j < Q; ++j) {
                    if (PAR == 0) StoreT<S>::st(a.f + oppq(j) * N + i, gc[j]);
                    else StoreT<S>::st(a.f + j * N + i + plain_off(j, SY, SZ), gc[j]);
                }
            } else {
                const Offs o(g, x, y, z);
#pragma unroll
                for (int j = 1; j < Q; ++j) {
                    const int32_t yj = i + int32_t(o.d(j));
                    if (PAR == 0) StoreT<S>::st(a.f + oppq(j) * N + i, gc[j]);
                    else StoreT<S>::st(a.f + j * N + yj, gc[j]);
                    const uint8_t ty = a.type[yj] & T_BASE;
                    if (ty == T_FLUID || ty == T_IF) continue;
                    const C bc = boundary_value<S, DIAG>(a, j, yj, ty, gc[j], n, sx, sy, sz, usq, mi, sPart);
                    if (PAR == 0) StoreT<S>::st(a.f + j * N + yj, bc);
                    else StoreT<S>::st(a.f + oppq(j) * N + i, bc);
                }
            }
            if (FSURF) {
                if ((t & T_BASE) == T_FLUID) {
                    a.mass[i] = M(n);
                } else {
                    a.aux[i] = n;
                    a.aux[N + i] = sx; a.aux[2 * N + i] = sy; a.aux[3 * N + i] = sz;
                    a.aux[4 * N + i] = hx; a.aux[5 * N + i] = hy; a.aux[6 * N + i] = hz;
                }
            }
        }
        // interface list append (wave-aggregated)
        if (FSURF) {
            const bool isIf = act && (t & T_BASE) == T_IF;
            const unsigned long long m = __ballot(isIf);
            if (m) {
                int base = 0;
                const int leader = __ffsll((unsigned long long)m) - 1;
                if (lane == leader) base = atomicAdd(&a.sc->nIf, __popcll(m));
                base = __shfl(base, leader, 64);
                if (isIf) {
                    const int pos = base + __popcll(m & ((1ull << lane) - 1ull));
                    if (pos < a.ifCap) a.ifList[pos] = i;
                    else atomicOr(&a.sc->listOverflow, 1);
                }
            }
        }
        // particle accumulators: wave segmented sums by particle, one LDS update per run
        const unsigned long long pm = __ballot(hkeyv >= 0);
        if (pm) {
            const int st0 = run_start(hkeyv, lane);
            const C s0 = seg_sum(dfx, st0, lane), s1 = seg_sum(dfy, st0, lane), s2 = seg_sum(dfz, st0, lane);
            const C s3 = seg_sum(dmx, st0, lane), s4 = seg_sum(dmy, st0, lane), s5 = seg_sum(dmz, st0, lane);
            const C s6 = seg_sum(dvol, st0, lane);
            const int nextKey = __shfl_down(hkeyv, 1, 64);
            const bool tail = hkeyv >= 0 && (lane == 63 || nextKey != hkeyv);
            if (tail) {
                int slot = hkeyv & (HS - 1);
                int probes = 0;
                for (; probes < HS; ++probes) {
                    const int old = atomicCAS(&hkey[slot], -1, hkeyv);
                    if (old == -1 || old == hkeyv) break;
                    slot = (slot + 1) & (HS - 1);
                }
                if (probes < HS) {
                    atomicAdd(&hval[slot][0], double(s0)); atomicAdd(&hval[slot][1], double(s1));
                    atomicAdd(&hval[slot][2], double(s2)); atomicAdd(&hval[slot][3], double(s3));
                    atomicAdd(&hval[slot][4], double(s4)); atomicAdd(&hval[slot][5], double(s5));
                    atomicAdd(&hval[slot][6], double(s6));
                } else {
                    double* hp = a.hacc + int64_t(hkeyv) * 8;
                    atomicAdd(hp + 0, double(s0)); atomicAdd(hp + 1, double(s1)); atomicAdd(hp + 2, double(s2));
                    atomicAdd(hp + 3, double(s3)); atomicAdd(hp + 4, double(s4)); atomicAdd(hp + 5, double(s5));
                    atomicAdd(hp + 6, double(s6));
                }
            }
        }
        // advance the column
        actc = actn;
        ic = in;
        tc = tn;
#pragma unroll
        for (int q = 0; q < Q; ++q) gc[q] = gn[q];
    }
    __syncthreads();
    for (int s = tid; s < HS; s += FX * FY) {
        const int key = hkey[s];
        if (key >= 0) {
            double* hp = a.hacc + int64_t(key) * 8;
            for (int c = 0; c < 7; ++c) atomicAdd(hp + c, hval[s][c]);
        }
    }
    if (tid < PART_STRIDE) a.blockPart[int64_t(tid) * gridDim.x + b0] = sPart[tid];
}

// edge bit of active nodes: a neighbour (wrapped) is not active, or the node sits next to a
// periodic face (its links wrap)
__global__ void k_mark_edges(Geom g, uint8_t* type) {
    const int32_t i = int32_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= int32_t(g.N)) return;
    const uint8_t t = type[i];
    if (!is_active(t)) return;
    int x, y, z;
    cell_xyz(g, i, x, y, z);
    bool edge = (x == 1 && g.per[0]) || (x == g.X - 2 && g.per[1]) || (y == 1 && g.per[2]) ||
                (y == g.Y - 2 && g.per[3]) || (z == 1 && g.per[4]) || (z == g.Z - 2 && g.per[5]);
    if (!edge) {
        const int32_t sy = g.X, sz = g.X * g.Y;
#pragma unroll
        for (int q = 1; q < Q; ++q) edge |= !is_active(type[i + plain_off(q, sy, sz)]);
    }
    const uint8_t nt = edge ? (t | T_EDGE) : (t & uint8_t(~T_EDGE));
    if (nt != t) type[i] = nt;
}

void lb_mark_edges(const Box& b, hipStream_t s) {
    hipLaunchKernelGGL(k_mark_edges, int((b.g.N + 255) / 256), 256, 0, s, b.g, b.type);
}

template <typename S>
static void launch_fused(Box& b, const Consts& k, const PartLB* plb, const ObjLB* olb, DevScalars* sc, bool diag,
                         hipStream_t s) {
    FusedArgs<S> a;
    a.g = b.g; a.f = (S*)b.f; a.type = b.type; a.solid = b.solid;
    a.mass = (typename Types<S>::M*)b.mass; a.aux = (typename Types<S>::C*)b.aux;
    a.ifList = b.ifList; a.ifCap = b.ifCap; a.plb = plb; a.olb = olb; a.hacc = b.hacc;
    a.blockPart = b.blockPart; a.sc = sc; a.k = k;
    a.nbx = (b.g.X + FX - 1) / FX; a.nby = (b.g.Y + FY - 1) / FY; a.nbz = (b.g.Z + FZ - 1) / FZ;
    const int nblk = a.nbx * a.nby * a.nbz;
    const int grid = ((nblk + 7) / 8) * 8;
    dim3 bs(FX, FY, 1);
    const bool fs = k.freeSurface != 0;
#define LBDEM_FUSED(P, F, D) hipLaunchKernelGGL((k_fused<S, P, F, D>), dim3(grid), bs, 0, s, a)
    if (b.parity == 0) {
        if (fs) { if (diag) LBDEM_FUSED(0, true, true); else LBDEM_FUSED(0, true, false); }
        else { if (diag) LBDEM_FUSED(0, false, true); else LBDEM_FUSED(0, false, false); }
    } else {
        if (fs) { if (diag) LBDEM_FUSED(1, true, true); else LBDEM_FUSED(1, true, false); }
        else { if (diag) LBDEM_FUSED(1, false, true); else LBDEM_FUSED(1, false, false); }
    }
#undef LBDEM_FUSED
    b.parity ^= 1;
}

int fused_grid_blocks(const Geom& g) {
    const int nbx = (g.X + FX - 1) / FX, nby = (g.Y + FY - 1) / FY, nbz = (g.Z + FZ - 1) / FZ;
    return ((nbx * nby * nbz + 7) / 8) * 8;
}

__global__ void k_begin_lb(DevScalars* sc, double* hacc, int nPart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) sc->nIf = 0;
    if (i < nPart * 8) hacc[i] = 0.0;
}

void lb_fused_step(Box& b, const Consts& k, const PartLB* plb, const ObjLB* olb, DevScalars* sc, bool diag,
                   hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    const int n = k.nPart * 8 + 1;
    hipLaunchKernelGGL(k_begin_lb, (n + 255) / 256, 256, 0, s, sc, b.hacc, k.nPart);
    if (ev0) (void)hipEventRecord(ev0, s);
    if (b.prec == LBDEM_PREC_FP64) launch_fused<double>(b, k, plb, olb, sc, diag, s);
    else if (b.prec == LBDEM_PREC_FP32) launch_fused<float>(b, k, plb, olb, sc, diag, s);
    else launch_fused<__half>(b, k, plb, olb, sc, diag, s);
    if (ev1) (void)hipEventRecord(ev1, s);
}

// deterministic sum of the per-block partials, unit conversion of all hydrodynamic results
// (LB.cpp:1966-1977, 1536-1543)
__global__ void k_hydro_finalize(int nblk, const double* part, Consts k, const double* hacc, double* FHydro,
                                 double* MHydro, double* fluidVol, DevScalars* sc, double* objFH, double* objMH,
                                 double* wallFH, int diag) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    if (blockIdx.x < PART_STRIDE) {
        // column c of the partial table (one block per column), summed in a fixed order
        const int c = blockIdx.x;
        if (c >= 7 * MAX_OBJ_PART && !diag) return;
        double acc = 0.0;
        for (int b = tid; b < nblk; b += blockDim.x) acc += part[int64_t(c) * nblk + b];
        red[tid] = acc;
        __syncthreads();
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if (tid < w) red[tid] += red[tid + w];
            __syncthreads();
        }
        if (tid == 0) {
            const double v = red[0];
            if (c < 7 * MAX_OBJ_PART) {
                const int o = c / 7, comp = c % 7;
                if (o < k.nObj) {
                    if (comp < 3) objFH[3 * o + comp] = v * k.unitForce;
                    else if (comp < 6) objMH[3 * o + comp - 3] = v * k.unitTorque;
                }
                if (c == 6) { sc->extraMass = v; sc->lastExtraMass = v; }
            } else {
                const int w = (c - 7 * MAX_OBJ_PART) / 3, comp = (c - 7 * MAX_OBJ_PART) % 3;
                if (w < k.nWall) wallFH[3 * w + comp] = v * k.unitForce;
            }
        }
    } else {
        const int p = (blockIdx.x - PART_STRIDE) * blockDim.x + tid;
        if (p < k.nPart) {
            const double* h = hacc + int64_t(p) * 8;
            for (int c = 0; c < 3; ++c) {
                FHydro[3 * p + c] = h[c] * k.unitForce;
                MHydro[3 * p + c] = h[3 + c] * k.unitTorque;
            }
            fluidVol[p] = h[6] * k.unitVolume;
        }
    }
}

void lb_hydro_finalize(const Box& b, const Consts& k, double* FHydro, double* MHydro, double* fluidVol,
                       DevScalars* sc, double* objFH, double* objMH, double* wallFH, bool diag, hipStream_t s) {
    const int nblk = fused_grid_blocks(b.g);
    const int grid = PART_STRIDE + (k.nPart + 255) / 256;
    hipLaunchKernelGGL(k_hydro_finalize, grid, 256, 0, s, nblk, b.blockPart, k, b.hacc, FHydro, MHydro, fluidVol,
                       sc, objFH, objMH, wallFH, diag ? 1 : 0);
}

// ------------------------------------------------------------------------------------------
// free surface (LB.cpp:1548-1852)
// ------------------------------------------------------------------------------------------
constexpr int LIST_GRID = 1024;

template <typename S, int L>
__global__ void k_fs_mass(Geom g, const S* f, const uint8_t* type, const typename Types<S>::M* mass,
                          const typename Types<S>::C* aux, const int32_t* ifList, typename Types<S>::M* newMass,
                          DevScalars* sc, int32_t* filledFlag) {
    using C = typename Types<S>::C;
    using M = typename Types<S>::M;
    const int nIf = sc->nIf;
    const double addE = sc->extraMass / double(nIf);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nIf; k += gridDim.x * blockDim.x) {
        const int64_t i = ifList[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        const C mi = C(double(mass[i]) + addE);
        C dm = 0;
        for (int j = 1; j < Q; ++j) {
            const int64_t yj = i + o.d(j);
            const uint8_t ty = type[yj] & T_BASE;
            if (ty != T_IF && ty != T_FLUID && ty != T_OBJ) continue;
            const C a1 = StoreT<S>::ld(f + int64_t(j) * g.N + yj);
            const C a2 = StoreT<S>::ld(f + int64_t(oppq(j)) * g.N + i);
            const C ms = L == 1 ? (a1 - a2) : (a2 - a1);  // incoming(opp j) - outgoing(j), offsets
            if (ty == T_IF) dm += C(0.5) * (C(double(mass[yj]) + addE) + mi) * ms;
            else if (ty == T_FLUID) dm += ms;
            else dm += mi * ms;
        }
        newMass[k] = M(mi + dm);
    }
}

__device__ inline void list_push(int32_t* list, int* count, int cap, int32_t v, int* overflow) {
    const int pos = atomicAdd(count, 1);
    if (pos < cap) list[pos] = v;
    else atomicOr(overflow, 1);
}

// assign new mass, classify filled / emptied (findInterfaceMutants, LB.cpp:1669-1698)
template <typename S>
__global__ void k_fs_mutants(uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux, int64_t N,
                             const int32_t* ifList, const typename Types<S>::M* newMass, int32_t* filled,
                             int32_t* emptied, int cap, DevScalars* sc, int32_t* marker) {
    const int nIf = sc->nIf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nIf; k += gridDim.x * blockDim.x) {
        const int64_t i = ifList[k];
        const auto m = newMass[k];
        mass[i] = m;
        const auto n = aux[i];
        if (m > n) {
            type[i] = (type[i] & T_PMASK) | T_FLUID;
            marker[i] = -3;
            list_push(filled, &sc->nFilled, cap, int32_t(i), &sc->listOverflow);
        } else if (m < 0) {
            type[i] = (type[i] & T_PMASK) | T_GAS;
            list_push(emptied, &sc->nEmptied, cap, int32_t(i), &sc->listOverflow);
        }
    }
}

// gas neighbours of filled nodes: donor = highest-index filled neighbour (LB.cpp:1704-1728)
__global__ void k_fs_donors(Geom g, const uint8_t* type, const int32_t* filled, int32_t* donor, int32_t* newIf,
                            int cap, DevScalars* sc) {
    const int nF = sc->nFilled;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nF; k += gridDim.x * blockDim.x) {
        const int64_t i = filled[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        for (int j = 1; j < Q; ++j) {
            const int64_t yj = i + o.d(j);
            if ((type[yj] & T_BASE) != T_GAS) continue;
            const int old = atomicMax(&donor[yj], int32_t(i));
            if (old < 0) list_push(newIf, &sc->nNewIf, cap, int32_t(yj), &sc->listOverflow);
        }
    }
}

// initialise the new interface nodes from their donor (node::initialize with the donor's
// u, visc and hydroForce+lbF, mass 0.01*rho0)
template <typename S, int L>
__global__ void k_fs_new_from_donor(Geom g, S* f, uint8_t* type, typename Types<S>::M* mass,
                                    typename Types<S>::C* aux, const int32_t* newIf, const int32_t* donor,
                                    DevScalars* sc, Consts k) {
    using C = typename Types<S>::C;
    const int nN = sc->nNewIf;
    for (int kk = blockIdx.x * blockDim.x + threadIdx.x; kk < nN; kk += gridDim.x * blockDim.x) {
        const int64_t gi = newIf[kk];
        const int64_t d = donor[gi];
        const int64_t N = g.N;
        type[gi] = (type[gi] & T_PMASK) | T_IF;
        const double n = k.rho0;
        const double ux = (double(aux[4 * N + d]) + k.lbF[0]) / 2.0 / n + double(aux[N + d]);
        const double uy = (double(aux[5 * N + d]) + k.lbF[1]) / 2.0 / n + double(aux[2 * N + d]);
        const double uz = (double(aux[6 * N + d]) + k.lbF[2]) / 2.0 / n + double(aux[3 * N + d]);
        int x, y, z;
        cell_xyz(g, gi, x, y, z);
        const Offs o(g, x, y, z);
        const double usq = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < Q; ++q) {
            const double vu = ux * cxq(q) + uy * cyq(q) + uz * czq(q);
            const double feq = wq(q) * n * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
            StoreT<S>::st(f + load_slot(g, gi, o, q, L), C(feq - wq(q)));
        }
        mass[gi] = typename Types<S>::M(0.01 * k.rho0);
        aux[gi] = C(n);
        aux[N + gi] = C(ux); aux[2 * N + gi] = C(uy); aux[3 * N + gi] = C(uz);
        aux[4 * N + gi] = C(0); aux[5 * N + gi] = C(0); aux[6 * N + gi] = C(0);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&sc->surplus, -0.01 * k.rho0 * double(nN));
}

// fluid neighbours of emptied nodes become interface with 99% of their density (LB.cpp:1730-1748)
template <typename S>
__global__ void k_fs_from_emptied(Geom g, uint8_t* type, typename Types<S>::M* mass, typename Types<S>::C* aux,
                                  const int32_t* emptied, int32_t* marker, int32_t* newIf, int cap, DevScalars* sc) {
    using C = typename Types<S>::C;
    using M = typename Types<S>::M;
    const int nE = sc->nEmptied;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nE; k += gridDim.x * blockDim.x) {
        const int64_t i = emptied[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        for (int j = 0; j < Q; ++j) {
            const int64_t l = i + o.d(j);
            if ((type[l] & T_BASE) != T_FLUID) continue;
            double nl;
            int old = atomicCAS(&marker[l], -1, -2);
            if (old == -1) {
                nl = double(mass[l]);               // ordinary fluid node: mass holds its density
            } else if (old == -3 && atomicCAS(&marker[l], -3, -4) == -3) {
                nl = double(aux[l]);                // filled this step: density kept in aux
            } else {
                continue;
            }
            type[l] = (type[l] & T_PMASK) | T_IF;
            mass[l] = M(0.99 * nl);
            aux[l] = C(nl);
            atomicAdd(&sc->surplus, 0.01 * nl);
            list_push(newIf, &sc->nNewIf, cap, int32_t(l), &sc->listOverflow);
        }
    }
}

// updateMutants (LB.cpp:1751-1790)
template <typename S>
__global__ void k_fs_update_mutants(uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux,
                                    const int32_t* filled, const int32_t* emptied, DevScalars* sc) {
    const int nF = sc->nFilled, nE = sc->nEmptied;
    double acc = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nF + nE; k += gridDim.x * blockDim.x) {
        if (k < nF) {
            const int64_t i = filled[k];
            if ((type[i] & T_BASE) == T_FLUID) {
                acc += double(mass[i]) - double(aux[i]);
                mass[i] = typename Types<S>::M(aux[i]);
            }
        } else {
            const int64_t i = emptied[k - nF];
            if ((type[i] & T_BASE) == T_GAS) acc += double(mass[i]);
        }
    }
    if (acc != 0.0) atomicAdd(&sc->surplus, acc);
}

// interface list after conversions: untouched old interface nodes + new ones
__global__ void k_fs_list2(const uint8_t* type, const int32_t* ifList, const int32_t* newIf, const int32_t* marker,
                           int32_t* list2, int cap, DevScalars* sc, int* count2) {
    const int nIf = sc->nIf, nN = sc->nNewIf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nIf + nN; k += gridDim.x * blockDim.x) {
        int64_t i;
        if (k < nIf) {
            i = ifList[k];
            if ((type[i] & T_BASE) != T_IF || marker[i] != -1) continue;
        } else {
            i = newIf[k - nIf];
        }
        list_push(list2, count2, cap, int32_t(i), &sc->listOverflow);
    }
}

// removeIsolated part 1: interface nodes without gas neighbours become fluid (LB.cpp:1796-1816)
template <typename S>
__global__ void k_fs_isolated_a(Geom g, uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux,
                                const int32_t* list2, const int* count2, DevScalars* sc) {
    const int n2 = *count2;
    double acc = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += gridDim.x * blockDim.x) {
        const int64_t i = list2[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        bool gasNb = false;
        for (int j = 1; j < Q && !gasNb; ++j) gasNb = (type[i + o.d(j)] & T_BASE) == T_GAS;
        if (!gasNb) {
            acc += double(mass[i]) - double(aux[i]);
            mass[i] = typename Types<S>::M(aux[i]);
            type[i] = (type[i] & T_PMASK) | T_FLUID;
        }
    }
    if (acc != 0.0) atomicAdd(&sc->surplus, acc);
}

// removeIsolated part 2: interface nodes without fluid neighbours become gas (LB.cpp:1821-1841)
template <typename S>
__global__ void k_fs_isolated_b(Geom g, uint8_t* type, const typename Types<S>::M* mass, const int32_t* list2,
                                const int* count2, DevScalars* sc) {
    const int n2 = *count2;
    double acc = 0.0;
    int keep = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += gridDim.x * blockDim.x) {
        const int64_t i = list2[k];
        if ((type[i] & T_BASE) != T_IF) continue;
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        bool fluidNb = false;
        for (int j = 1; j < Q && !fluidNb; ++j) fluidNb = (type[i + o.d(j)] & T_BASE) == T_FLUID;
        if (!fluidNb) {
            acc += double(mass[i]);
            type[i] = (type[i] & T_PMASK) | T_GAS;
        } else {
            keep++;
        }
    }
    if (acc != 0.0) atomicAdd(&sc->surplus, acc);
    if (keep) atomicAdd(&sc->nIfFinal, keep);
}

// redistributeMass (LB.cpp:1844-1852) and marker reset
template <typename S>
__global__ void k_fs_redistribute(const uint8_t* type, typename Types<S>::M* mass, const int32_t* list2,
                                  const int* count2, int32_t* marker, const int32_t* filled, const int32_t* newIf,
                                  DevScalars* sc) {
    const int n2 = *count2;
    const double add = sc->surplus / double(sc->nIfFinal);
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += gridDim.x * blockDim.x) {
        const int64_t i = list2[k];
        if ((type[i] & T_BASE) == T_IF) mass[i] = typename Types<S>::M(double(mass[i]) + add);
    }
    const int nF = sc->nFilled, nN = sc->nNewIf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nF + nN; k += gridDim.x * blockDim.x)
        marker[k < nF ? filled[k] : newIf[k - nF]] = -1;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->lastSurplus = sc->surplus;
}

__global__ void k_fs_begin(DevScalars* sc, int* count2) {
    sc->surplus = 0.0;
    sc->nFilled = sc->nEmptied = sc->nNewIf = 0;
    sc->nIfFinal = 0;
    *count2 = 0;
}

template <typename S>
static void fs_impl(Box& b, const Consts& k, DevScalars* sc, hipStream_t s) {
    using M = typename Types<S>::M;
    using C = typename Types<S>::C;
    S* f = (S*)b.f;
    M* mass = (M*)b.mass;
    C* aux = (C*)b.aux;
    int* count2 = (int*)(b.ifList2 + b.ifCap);  // counter stored after the list
    const int grid = LIST_GRID, bs = 256;
    hipLaunchKernelGGL(k_fs_begin, 1, 1, 0, s, sc, count2);
    if (b.parity == 1)
        hipLaunchKernelGGL((k_fs_mass<S, 1>), grid, bs, 0, s, b.g, f, b.type, mass, aux, b.ifList, (M*)b.ifNewMass, sc, b.donor);
    else
        hipLaunchKernelGGL((k_fs_mass<S, 0>), grid, bs, 0, s, b.g, f, b.type, mass, aux, b.ifList, (M*)b.ifNewMass, sc, b.donor);
    hipLaunchKernelGGL(k_fs_mutants<S>, grid, bs, 0, s, b.type, mass, aux, b.g.N, b.ifList, (M*)b.ifNewMass, b.filled,
                       b.emptied, b.ifCap, sc, b.donor);
    hipLaunchKernelGGL(k_fs_donors, grid, bs, 0, s, b.g, b.type, b.filled, b.donor, b.newIf, b.ifCap, sc);
    if (b.parity == 1)
        hipLaunchKernelGGL((k_fs_new_from_donor<S, 1>), grid, bs, 0, s, b.g, f, b.type, mass, aux, b.newIf, b.donor, sc, k);
    else
        hipLaunchKernelGGL((k_fs_new_from_donor<S, 0>), grid, bs, 0, s, b.g, f, b.type, mass, aux, b.newIf, b.donor, sc, k);
    hipLaunchKernelGGL(k_fs_from_emptied<S>, grid, bs, 0, s, b.g, b.type, mass, aux, b.emptied, b.donor, b.newIf, b.ifCap, sc);
    hipLaunchKernelGGL(k_fs_update_mutants<S>, grid, bs, 0, s, b.type, mass, aux, b.filled, b.emptied, sc);
    hipLaunchKernelGGL(k_fs_list2, grid, bs, 0, s, b.type, b.ifList, b.newIf, b.donor, b.ifList2, b.ifCap, sc, count2);
    hipLaunchKernelGGL(k_fs_isolated_a<S>, grid, bs, 0, s, b.g, b.type, mass, aux, b.ifList2, count2, sc);
    hipLaunchKernelGGL(k_fs_isolated_b<S>, grid, bs, 0, s, b.g, b.type, mass, b.ifList2, count2, sc);
    hipLaunchKernelGGL(k_fs_redistribute<S>, grid, bs, 0, s, b.type, mass, b.ifList2, count2, b.donor, b.filled, b.newIf, sc);
}

void lb_free_surface(Box& b, const Consts& k, DevScalars* sc, hipStream_t s) {
    if (b.prec == LBDEM_PREC_FP64) fs_impl<double>(b, k, sc, s);
    else if (b.prec == LBDEM_PREC_FP32) fs_impl<float>(b, k, sc, s);
    else fs_impl<__half>(b, k, sc, s);
}

// ------------------------------------------------------------------------------------------
// particle flags (LB.cpp:683-711 via 1875-1899; findNewActive LB.cpp:1980-2029)
// ------------------------------------------------------------------------------------------
// one block per particle over its bounding box: active nodes inside the sphere, highest index wins
__global__ void k_flags_compute(Geom g, const uint8_t* type, const PartLB* plb, int nPart, int32_t* flagNew,
                                const DevScalars* sc) {
    if (!sc->newNeighborList) return;
    const int p = blockIdx.x;
    if (p >= nPart) return;
    const PartLB P = plb[p];
    int lo[3], hi[3];
    const int dims[3] = {g.X, g.Y, g.Z};
    for (int a = 0; a < 3; ++a) {
        lo[a] = int(floor(P.c[a] - P.r)) - 1;
        hi[a] = int(ceil(P.c[a] + P.r)) + 1;
        lo[a] = lo[a] < 0 ? 0 : lo[a];
        hi[a] = hi[a] > dims[a] - 1 ? dims[a] - 1 : hi[a];
    }
    if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) return;
    const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nz = hi[2] - lo[2] + 1;
    const int tot = nx * ny * nz;
    for (int c = threadIdx.x; c < tot; c += blockDim.x) {
        const int x = lo[0] + c % nx, y = lo[1] + (c / nx) % ny, z = lo[2] + c / (nx * ny);
        const int64_t i = cidx(g, x, y, z);
        if (!is_active(type[i])) continue;
        if (inside_sphere(x, y, z, P.c[0], P.c[1], P.c[2], P.r)) atomicMax(&flagNew[i], p);
    }
}

void lb_flags_rebuild_compute(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, hipStream_t s) {
    if (k.nPart > 0) hipLaunchKernelGGL(k_flags_compute, k.nPart, 256, 0, s, b.g, b.type, plb, k.nPart, b.flagNew, sc);
}

// clear every flag and solid index, apply the new flags (initializeParticleBoundaries)
__global__ void k_flags_apply(int64_t N, uint8_t* type, int32_t* solid, int32_t* flagNew, DevScalars* sc) {
    if (!sc->newNeighborList) return;
    __shared__ int bmin;
    if (threadIdx.x == 0) bmin = INT_MAX;
    __syncthreads();
    int mymin = INT_MAX;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
        const int32_t s = flagNew[i];
        if (s >= 0) {
            type[i] = (type[i] & T_BASE) | T_PMASK;
            solid[i] = s;
            flagNew[i] = -1;
            mymin = min(mymin, int(i));
        } else {
            type[i] &= T_BASE;
            solid[i] = 0;
        }
    }
    if (mymin != INT_MAX) atomicMin(&bmin, mymin);
    __syncthreads();
    if (threadIdx.x == 0 && bmin != INT_MAX) atomicMin(&sc->kminP, bmin);
}

__global__ void k_flags_apply_begin(DevScalars* sc) {
    if (sc->newNeighborList) sc->kminP = INT_MAX;
}
__global__ void k_flags_apply_end(DevScalars* sc) {
    if (!sc->newNeighborList) return;
    if (sc->kminP == INT_MAX) sc->kminP = -1;
    sc->newNeighborList = 0;
}

void lb_flags_rebuild_apply(const Box& b, DevScalars* sc, hipStream_t s) {
    hipLaunchKernelGGL(k_flags_apply_begin, 1, 1, 0, s, sc);
    const int64_t nb = (b.g.N + 255) / 256;
    const int grid = int(nb < 65536 ? nb : 65536);
    hipLaunchKernelGGL(k_flags_apply, grid, 256, 0, s, b.g.N, b.type, b.solid, b.flagNew, sc);
    hipLaunchKernelGGL(k_flags_apply_end, 1, 1, 0, s, sc);
}

__global__ void k_fna_check(Geom g, const int32_t* solid, const PartLB* plb, int nPart, DevScalars* sc) {
    const int kp = sc->kminP;
    sc->needScan = 0;
    if (kp < 0) return;
    int x, y, z;
    cell_xyz(g, kp, x, y, z);
    const int sp = solid[kp];
    if (sp < 0 || sp >= nPart) { sc->needScan = 1; sc->kstar = INT_MAX; return; }
    const PartLB P = plb[sp];
    if (!inside_sphere(x, y, z, P.c[0], P.c[1], P.c[2], P.r)) { sc->needScan = 1; sc->kstar = INT_MAX; }
}

__global__ void k_fna_scan(Geom g, const uint8_t* type, const int32_t* solid, const PartLB* plb, int nPart,
                           DevScalars* sc) {
    if (!sc->needScan) return;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        if (!(type[i] & T_PMASK)) continue;
        const int sp = solid[i];
        if (sp < 0 || sp >= nPart) continue;
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const PartLB P = plb[sp];
        if (inside_sphere(x, y, z, P.c[0], P.c[1], P.c[2], P.r)) atomicMin(&sc->kstar, int(i));
    }
}

__global__ void k_fna_unflag(Geom g, uint8_t* type, int32_t* solid, DevScalars* sc) {
    if (!sc->needScan) return;
    const int ks = sc->kstar;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N && i < ks; i += int64_t(gridDim.x) * blockDim.x) {
        if (type[i] & T_PMASK) { type[i] &= T_BASE; solid[i] = 0; }
    }
}
__global__ void k_fna_end(DevScalars* sc) {
    if (!sc->needScan) return;
    sc->kminP = sc->kstar == INT_MAX ? -1 : sc->kstar;
    sc->needScan = 0;
}

void lb_find_new_active(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, hipStream_t s) {
    hipLaunchKernelGGL(k_fna_check, 1, 1, 0, s, b.g, b.solid, plb, k.nPart, sc);
    const int64_t nb = (b.g.N + 255) / 256;
    const int grid = int(nb < 16384 ? nb : 16384);
    hipLaunchKernelGGL(k_fna_scan, grid, 256, 0, s, b.g, b.type, b.solid, plb, k.nPart, sc);
    hipLaunchKernelGGL(k_fna_unflag, grid, 256, 0, s, b.g, b.type, b.solid, sc);
    hipLaunchKernelGGL(k_fna_end, 1, 1, 0, s, sc);
}

// ------------------------------------------------------------------------------------------
// impactor nodes (updateObjectBoundaries + findNewActiveObject, LB.cpp:797-943)
// ------------------------------------------------------------------------------------------
// uncovered object nodes (LB.cpp:846-943). All threads of one block test the previous object
// nodes (ascending list) against the moved sphere and compact the uncovered ones in ascending
// order; one thread then converts them in that order, because each decision reads the node
// types its predecessors wrote.
constexpr int UNCOVER_THREADS = 1024;
template <typename S, int L>
__global__ __launch_bounds__(UNCOVER_THREADS) void k_obj_uncover(Geom g, S* f, uint8_t* type,
                                                                 typename Types<S>::M* mass,
                                                                 typename Types<S>::C* aux, const int32_t* objList,
                                                                 int32_t* cand, DevScalars* sc, const ObjLB* olb,
                                                                 Consts k) {
    using C = typename Types<S>::C;
    __shared__ int counts[UNCOVER_THREADS];
    __shared__ int total;
    const int n = sc->nObjList;
    const ObjLB O = olb[0];
    const int per = (n + blockDim.x - 1) / blockDim.x;
    const int beg = threadIdx.x * per, end = min(n, beg + per);
    int cnt = 0;
    for (int kk = beg; kk < end; ++kk) {
        int x, y, z;
        cell_xyz(g, objList[kk], x, y, z);
        cnt += !inside_sphere(x, y, z, O.c[0], O.c[1], O.c[2], O.r);
    }
    counts[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < int(blockDim.x); ++t) { const int c = counts[t]; counts[t] = acc; acc += c; }
        total = acc;
    }
    __syncthreads();
    int pos = counts[threadIdx.x];
    for (int kk = beg; kk < end; ++kk) {
        const int32_t i = objList[kk];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        if (!inside_sphere(x, y, z, O.c[0], O.c[1], O.c[2], O.r)) cand[pos++] = i;
    }
    __threadfence_block();
    __syncthreads();
    if (threadIdx.x != 0) return;
    volatile uint8_t* vt = type;
    const int m = total;
    for (int kk = 0; kk < m; ++kk) {
        const int64_t i = cand[kk];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        bool anyFluid = false;
        for (int j = 1; j < Q; ++j) anyFluid |= (vt[nbr_plain(g, x, y, z, j)] & T_BASE) == T_FLUID;
        const uint8_t pbit = vt[i] & T_PMASK;
        if (!anyFluid) { vt[i] = pbit | T_GAS; continue; }
        bool anyGas = false;
        for (int j = 1; j < Q; ++j) anyGas |= (vt[nbr_plain(g, x, y, z, j)] & T_BASE) == T_GAS;
        const bool isIf = anyGas;
        vt[i] = pbit | (isIf ? T_IF : T_FLUID);
        const double nn = k.rho0;
        const double ux = k.lbF[0] / 2.0 / nn + (isIf ? 0.0 : k.initVel[0]);
        const double uy = k.lbF[1] / 2.0 / nn + (isIf ? 0.0 : k.initVel[1]);
        const double uz = k.lbF[2] / 2.0 / nn + (isIf ? 0.0 : k.initVel[2]);
        const double usq = ux * ux + uy * uy + uz * uz;
        const Offs o(g, x, y, z);
        for (int q = 0; q < Q; ++q) {
            const double vu = ux * cxq(q) + uy * cyq(q) + uz * czq(q);
            const double feq = wq(q) * nn * (1.0 + 3.0 * vu + 4.5 * vu * vu - 1.5 * usq);
            StoreT<S>::st(f + load_slot(g, i, o, q, L), C(feq - wq(q)));
        }
        mass[i] = typename Types<S>::M(isIf ? 0.5 * k.rho0 : k.rho0);
        aux[i] = C(nn);
        aux[g.N + i] = C(ux); aux[2 * g.N + i] = C(uy); aux[3 * g.N + i] = C(uz);
        aux[4 * g.N + i] = C(0); aux[5 * g.N + i] = C(0); aux[6 * g.N + i] = C(0);
    }
}

// cover: every node inside an object sphere becomes an object node; ordered object node list
__global__ void k_obj_cover(Geom g, uint8_t* type, int32_t* solid, const ObjLB* olb, int nObj, int32_t* objList, int cap,
                            DevScalars* sc) {
    __shared__ int counts[1024];
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    const int dims[3] = {g.X, g.Y, g.Z};
    for (int o = 0; o < nObj; ++o)
        for (int a = 0; a < 3; ++a) {
            const int l = int(floor(olb[o].c[a] - olb[o].r)) - 1, h = int(ceil(olb[o].c[a] + olb[o].r)) + 1;
            lo[a] = min(lo[a], max(l, 0));
            hi[a] = max(hi[a], min(h, dims[a] - 1));
        }
    const int lo0 = lo[0], lo1 = lo[1], lo2 = lo[2];
    const int nx = max(hi[0] - lo0 + 1, 0), ny = max(hi[1] - lo1 + 1, 0), nz = max(hi[2] - lo2 + 1, 0);
    const int tot = nx * ny * nz;
    const int per = (tot + blockDim.x - 1) / blockDim.x;
    const int beg = threadIdx.x * per, end = min(tot, beg + per);
    int cnt = 0;
    for (int c = beg; c < end; ++c) {
        const int x = lo0 + c % nx, y = lo1 + (c / nx) % ny, z = lo2 + c / (nx * ny);
        int which = -1;
        for (int o = 0; o < nObj; ++o)
            if (inside_sphere(x, y, z, olb[o].c[0], olb[o].c[1], olb[o].c[2], olb[o].r)) which = o;
        if (which >= 0) {
            const int64_t i = cidx(g, x, y, z);
            type[i] = (type[i] & T_PMASK) | T_OBJ;
            solid[i] = which;
            cnt++;
        }
    }
    counts[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < int(blockDim.x); ++t) { const int c = counts[t]; counts[t] = acc; acc += c; }
        sc->nObjList = acc < cap ? acc : cap;
        if (acc > cap) sc->listOverflow |= 1;
    }
    __syncthreads();
    int pos = counts[threadIdx.x];
    for (int c = beg; c < end; ++c) {
        const int x = lo0 + c % nx, y = lo1 + (c / nx) % ny, z = lo2 + c / (nx * ny);
        const int64_t i = cidx(g, x, y, z);
        if ((type[i] & T_BASE) == T_OBJ) {
            bool in = false;
            for (int o = 0; o < nObj && !in; ++o)
                in = inside_sphere(x, y, z, olb[o].c[0], olb[o].c[1], olb[o].c[2], olb[o].r);
            if (in) { if (pos < cap) objList[pos] = int32_t(i); pos++; }
        }
    }
}

template <typename S>
static void obj_impl(Box& b, const Consts& k, const ObjLB* olb, DevScalars* sc, hipStream_t s) {
    using M = typename Types<S>::M;
    using C = typename Types<S>::C;
    if (b.parity == 0)
        hipLaunchKernelGGL((k_obj_uncover<S, 0>), 1, UNCOVER_THREADS, 0, s, b.g, (S*)b.f, b.type, (M*)b.mass, (C*)b.aux,
                           b.objList, b.objCand, sc, olb, k);
    else
        hipLaunchKernelGGL((k_obj_uncover<S, 1>), 1, UNCOVER_THREADS, 0, s, b.g, (S*)b.f, b.type, (M*)b.mass, (C*)b.aux,
                           b.objList, b.objCand, sc, olb, k);
    hipLaunchKernelGGL(k_obj_cover, 1, 1024, 0, s, b.g, b.type, b.solid, olb, k.nObj, b.objList, b.objCap, sc);
}

void lb_object_update(Box& b, const Consts& k, const ObjLB* olb, DevScalars* sc, hipStream_t s) {
    if (k.nObj <= 0) return;
    if (b.prec == LBDEM_PREC_FP64) obj_impl<double>(b, k, olb, sc, s);
    else if (b.prec == LBDEM_PREC_FP32) obj_impl<float>(b, k, olb, sc, s);
    else obj_impl<__half>(b, k, olb, sc, s);
}

// ------------------------------------------------------------------------------------------
// downloads
// ------------------------------------------------------------------------------------------
template <typename S>
__global__ void k_download_f(Geom g, const S* f, const uint8_t* type, int L, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        const bool act = is_active(type[i]);
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        if (!act) { for (int q = 0; q < Q; ++q) out[int64_t(q) * g.N + i] = 0.0; continue; }
        const Offs o(g, x, y, z);
        for (int q = 0; q < Q; ++q)
            out[int64_t(q) * g.N + i] = wq(q) + double(StoreT<S>::ld(f + load_slot(g, i, o, q, L)));
    }
}
void lb_download_f(const Box& b, double* out, hipStream_t s) {
    const int grid = 4096;
    if (b.prec == LBDEM_PREC_FP64) hipLaunchKernelGGL(k_download_f<double>, grid, 256, 0, s, b.g, (const double*)b.f, b.type, b.parity, out);
    else if (b.prec == LBDEM_PREC_FP32) hipLaunchKernelGGL(k_download_f<float>, grid, 256, 0, s, b.g, (const float*)b.f, b.type, b.parity, out);
    else hipLaunchKernelGGL(k_download_f<__half>, grid, 256, 0, s, b.g, (const __half*)b.f, b.type, b.parity, out);
}

template <typename S>
__global__ void k_download_n(Geom g, const uint8_t* type, const typename Types<S>::M* mass,
                             const typename Types<S>::C* aux, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        const uint8_t t = type[i] & T_BASE;
        out[i] = t == T_FLUID ? double(mass[i]) : (t == T_IF ? double(aux[i]) : 0.0);
    }
}
void lb_download_n(const Box& b, const Consts& k, double* out, hipStream_t s) {
    (void)k;
    if (b.prec == LBDEM_PREC_FP64) hipLaunchKernelGGL(k_download_n<double>, 4096, 256, 0, s, b.g, b.type, (const double*)b.mass, (const double*)b.aux, out);
    else hipLaunchKernelGGL(k_download_n<float>, 4096, 256, 0, s, b.g, b.type, (const float*)b.mass, (const float*)b.aux, out);
}

// total fluid mass: active nodes outside particles (LB.cpp:388-395)
template <typename M>
__global__ void k_total_mass(int64_t N, const uint8_t* type, const M* mass, double* out) {
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
        const uint8_t t = type[i];
        if (!(t & T_PMASK) && ((t & T_BASE) == T_FLUID || (t & T_BASE) == T_IF)) acc += double(mass[i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}
void lb_total_mass(const Box& b, double* out, hipStream_t s) {
    // out: 1024 per-block partials, summed on the host in block order
    if (b.prec == LBDEM_PREC_FP64) hipLaunchKernelGGL(k_total_mass<double>, 1024, 256, 0, s, b.g.N, b.type, (const double*)b.mass, out);
    else hipLaunchKernelGGL(k_total_mass<float>, 1024, 256, 0, s, b.g.N, b.type, (const float*)b.mass, out);
}

__global__ void k_count_types(int64_t N, const uint8_t* type, int* counts) {
    int c[5] = {0, 0, 0, 0, 0};
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
        const uint8_t t = type[i];
        const uint8_t b = t & T_BASE;
        c[0] += b == T_FLUID; c[1] += b == T_IF; c[2] += b == T_OBJ; c[3] += (t & T_PMASK) != 0;
        c[4] += b == T_GAS;
    }
    for (int j = 0; j < 5; ++j) if (c[j]) atomicAdd(&counts[j], c[j]);
}
void lb_count_types(const Box& b, int* counts, hipStream_t s) {
    hipLaunchKernelGGL(k_count_types, 1024, 256, 0, s, b.g.N, b.type, counts);
}

}  // namespace lbdem
