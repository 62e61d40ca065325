// Lattice kernels of the coupled step on gfx950:
//   * k_fused      — reconstruct + particle interior forcing + BGK/Guo collision + AA streaming
//                    with gas / wall / moving-object boundary conditions resolved by the storing
//                    cell (reference LB.cpp:1261-1544, 1901-1978, node.cpp:75-207)
//   * free-surface mass exchange and interface conversion (LB.cpp:1548-1852)
//   * particle flags and the moving-object (impactor) node update (LB.cpp:683-711, 797-943,
//     1875-1899, 1980-2029; findNewSolid, LB.cpp:2032-2100, changes no flag)
#include "lbdem_internal.hpp"
#include "launch.hpp"

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
    if ((uint64_t(i) | uint64_t(g.X) * uint64_t(g.Y)) <= 0xffffffffull) {  // 32-bit quotients
        const uint32_t u = uint32_t(i), sxy = uint32_t(g.X) * uint32_t(g.Y);
        const uint32_t zz = u / sxy, rem = u - zz * sxy, yy = rem / uint32_t(g.X);
        z = int(zz);
        y = int(yy);
        x = int(rem - yy * uint32_t(g.X));
        return;
    }
    z = int(i / (int64_t(g.X) * g.Y));
    const int64_t rem = i - int64_t(z) * g.X * g.Y;
    y = int(rem / g.X);
    x = int(rem - int64_t(y) * g.X);
}

// lattice vectors for a direction index known only at run time: 2-bit fields of packed words
constexpr uint64_t pack_dir(int axis) {
    uint64_t k = 0;
    for (int q = 0; q < Q; ++q) k |= uint64_t((axis == 0 ? cxq(q) : axis == 1 ? cyq(q) : czq(q)) + 1) << (2 * q);
    return k;
}
constexpr uint64_t PACK_CX = pack_dir(0), PACK_CY = pack_dir(1), PACK_CZ = pack_dir(2);
constexpr bool opp_rule_ok() {
    for (int q = 1; q < Q; ++q)
        if (oppq(q) != ((q - 1) ^ 1) + 1) return false;
    return true;
}
static_assert(opp_rule_ok(), "opposite-direction rule");
__device__ inline int cx_rt(int q) { return int((PACK_CX >> (2 * q)) & 3u) - 1; }
__device__ inline int cy_rt(int q) { return int((PACK_CY >> (2 * q)) & 3u) - 1; }
__device__ inline int cz_rt(int q) { return int((PACK_CZ >> (2 * q)) & 3u) - 1; }
__device__ inline int opp_rt(int q) { return q == 0 ? 0 : ((q - 1) ^ 1) + 1; }
__device__ inline double w_rt(int q) { return q == 0 ? 12.0 / 36.0 : (q < 7 ? 2.0 / 36.0 : 1.0 / 36.0); }
// Per-axis index offsets of the wrapped neighbour rule: nb(q) = i + ox[cx+1] + oy[cy+1] + oz[cz+1]
// (offsets fit 32 bits: initialisation rejects boxes with 19*N >= 2^31)
struct Offs {
    int32_t ox[3], oy[3], oz[3];
    __device__ Offs(const Geom& g, int x, int y, int z) {
        const int32_t sy = g.X, sz = g.X * g.Y;
        ox[1] = 0; oy[1] = 0; oz[1] = 0;
        ox[0] = (x - 1 == 0 && g.per[0]) ? g.X - 3 : -1;
        ox[2] = (x + 1 == g.X - 1 && g.per[1]) ? -(g.X - 3) : 1;
        oy[0] = (y - 1 == 0 && g.per[2]) ? (g.Y - 3) * sy : -sy;
        oy[2] = (y + 1 == g.Y - 1 && g.per[3]) ? -(g.Y - 3) * sy : sy;
        oz[0] = (z - 1 == 0 && g.per[4]) ? (g.Z - 3) * sz : -sz;
        oz[2] = (z + 1 == g.Z - 1 && g.per[5]) ? -(g.Z - 3) * sz : sz;
    }
    // selects and packed direction words, not indexed arrays: a direction known only at run
    // time must not put tables or offsets in scratch memory
    __device__ int32_t d(int q) const {
        const int cx = cx_rt(q), cy = cy_rt(q), cz = cz_rt(q);
        return (cx < 0 ? ox[0] : (cx > 0 ? ox[2] : 0)) + (cy < 0 ? oy[0] : (cy > 0 ? oy[2] : 0)) +
               (cz < 0 ? oz[0] : (cz > 0 ? oz[2] : 0));
    }
};

__device__ inline int32_t offs_rt(const Offs& o, int q) { return o.d(q); }

// slot holding the post-streaming population q of cell i in layout L (0 natural, 1 swapped)
__device__ inline int64_t load_slot(const Geom& g, int64_t i, const Offs& o, int q, int L) {
    return L == 0 ? int64_t(q) * g.N + i : int64_t(opp_rt(q)) * g.N + i + o.d(opp_rt(q));
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
__device__ inline void k_init_fields(Geom g, S* f, const uint8_t* type, typename Types<S>::M* mass,
                              typename Types<S>::C* aux, typename Types<S>::C* visc, Consts k) {
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
    // density and velocity of every active node (the interface nodes' for the free surface, all
    // of them for a velocity read-back before the first LB step)
    aux[i] = C(n);
    aux[g.N + i] = C(ux); aux[2 * g.N + i] = C(uy); aux[3 * g.N + i] = C(uz);
    aux[4 * g.N + i] = C(0); aux[5 * g.N + i] = C(0); aux[6 * g.N + i] = C(0);
    if (visc) visc[i] = C(k.visc);  // initDynVisc (LB.cpp:1083-1090)
}

// restartVariables (LB.cpp:1098-1138) for record r of the restart file, at active node node[r]:
// populations f = restF, mass = restMass + 1, density n = restN + 1 and velocity u (node::restart,
// node.cpp:38-51, with initDensity 1); hydrodynamic force zero (a fresh node)
template <typename S>
__device__ inline void k_restart_fields(Geom g, S* f, typename Types<S>::M* mass, typename Types<S>::C* aux,
                                        typename Types<S>::C* visc, const double* rec, const int32_t* node, int nrec) {
    using C = typename Types<S>::C;
    const int r = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (r >= nrec) return;
    const int64_t i = node[r];
    const double* v = rec + int64_t(r) * LBDEM_RESTART_REC;
    for (int q = 0; q < Q; ++q) StoreT<S>::st(f + int64_t(q) * g.N + i, C(v[6 + q] - wq(q)));
    mass[i] = typename Types<S>::M(v[4] + 1.0);
    aux[i] = C(v[0] + 1.0);
    aux[g.N + i] = C(v[1]); aux[2 * g.N + i] = C(v[2]); aux[3 * g.N + i] = C(v[3]);
    aux[4 * g.N + i] = C(0); aux[5 * g.N + i] = C(0); aux[6 * g.N + i] = C(0);
    if (visc) visc[i] = C(v[5]);
}
void lb_restart_fields(const Box& b, const double* rec, const int32_t* node, int nrec, hipStream_t s) {
    const int bs = 256, nb = (nrec + bs - 1) / bs;
    if (b.prec == LBDEM_PREC_FP64)
        launch<&k_restart_fields<double>, 1024, 1>(nb, bs, s, b.g, (double*)b.f, (double*)b.mass, (double*)b.aux,
                                                   (double*)b.visc, rec, node, nrec);
    else if (b.prec == LBDEM_PREC_FP32)
        launch<&k_restart_fields<float>, 1024, 1>(nb, bs, s, b.g, (float*)b.f, (float*)b.mass, (float*)b.aux,
                                                  (float*)b.visc, rec, node, nrec);
    else
        launch<&k_restart_fields<__half>, 1024, 1>(nb, bs, s, b.g, (__half*)b.f, (float*)b.mass, (float*)b.aux,
                                                   (float*)b.visc, rec, node, nrec);
}

// node viscosities (the per-node plane), zero at nodes that are not active
template <typename C>
__device__ inline void k_download_visc(int64_t N, const uint8_t* type, const C* visc, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = is_active(type[i]) ? double(visc[i]) : 0.0;
}
void lb_download_visc(const Box& b, double* out, hipStream_t s) {
    if (b.prec == LBDEM_PREC_FP64) launch<&k_download_visc<double>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const double*)b.visc, out);
    else launch<&k_download_visc<float>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const float*)b.visc, out);
}

// velocity of every active node (aux planes 1-3), x y z per node, zero elsewhere
template <typename C>
__device__ inline void k_download_vel(int64_t N, const uint8_t* type, const C* aux, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
        const bool act = is_active(type[i]);
        for (int a = 0; a < 3; ++a) out[3 * i + a] = act ? double(aux[(1 + a) * N + i]) : 0.0;
    }
}
void lb_download_velocity(const Box& b, double* out, hipStream_t s) {
    if (b.prec == LBDEM_PREC_FP64) launch<&k_download_vel<double>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const double*)b.aux, out);
    else launch<&k_download_vel<float>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const float*)b.aux, out);
}

// SHydro of every particle (LB.cpp:1951-1959): the sum over the active nodes inside it of
// radius (x) diffVel in physical units, with diffVel = -(the node's hydrodynamic force), which the
// LB step just stored in aux planes 4-6 (it runs with every active node's values stored). One
// thread per (x, y) column and run of HS_ZRUN nodes in z walks its run and sums in registers while
// the particle is unchanged; the sums go to the block's LDS table (global atomics when it is
// full), which adds them to the particles once per block.
__device__ inline void k_zero_doubles(double* p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = 0.0;
}
constexpr int HS_ZRUN = 16, HS_SLOTS = 128;
__device__ inline void hydro_stress_add(int key, const double* acc, int* hk, double (*hv)[9], double* shydro) {
    int slot = key & (HS_SLOTS - 1), probes = 0;
    for (; probes < HS_SLOTS; ++probes) {
        const int old = atomicCAS(&hk[slot], -1, key);
        if (old == -1 || old == key) break;
        slot = (slot + 1) & (HS_SLOTS - 1);
    }
    if (probes < HS_SLOTS)
        for (int e = 0; e < 9; ++e) atomicAdd(&hv[slot][e], acc[e]);
    else
        for (int e = 0; e < 9; ++e) atomicAdd(shydro + 9 * int64_t(key) + e, acc[e]);
}
template <typename M>
__device__ inline void k_hydro_stress(Geom g, const uint8_t* type, const int32_t* solid, const M* aux,
                                      const PartLB* plb, double L, double F, double* shydro) {
#pragma clang fp contract(off)
    __shared__ int hk[HS_SLOTS];
    __shared__ double hv[HS_SLOTS][9];
    for (int s = threadIdx.x; s < HS_SLOTS; s += blockDim.x) {
        hk[s] = -1;
        for (int e = 0; e < 9; ++e) hv[s][e] = 0.0;
    }
    __syncthreads();
    const int64_t SZ = int64_t(g.X) * g.Y, N = g.N;
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t col = t % SZ, run = t / SZ;
    const int x = int(col % g.X), y = int(col / g.X);
    const int z0 = int(run) * HS_ZRUN, z1 = min(g.Z, z0 + HS_ZRUN);  // empty past the last run
    int key = -1;
    double acc[9];
    for (int e = 0; e < 9; ++e) acc[e] = 0.0;
    for (int z = z0; z < z1; ++z) {
        const int64_t i = col + SZ * z;
        const uint8_t ty = type[i];
        const int k = (is_active(ty) && (ty & T_PMASK)) ? solid[i] : -1;
        if (k != key) {
            if (key >= 0) hydro_stress_add(key, acc, hk, hv, shydro);
            key = k;
            for (int e = 0; e < 9; ++e) acc[e] = 0.0;
        }
        if (k < 0) continue;
        const PartLB P = plb[k];
        const double r[3] = {double(x) - P.c[0], double(y) - P.c[1], double(z) - P.c[2]};
        const double d[3] = {-double(aux[4 * N + i]), -double(aux[5 * N + i]), -double(aux[6 * N + i])};
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) acc[3 * a + c] += r[a] * L * d[c] * F;
    }
    if (key >= 0) hydro_stress_add(key, acc, hk, hv, shydro);
    __syncthreads();
    for (int s = threadIdx.x; s < HS_SLOTS; s += blockDim.x)
        if (hk[s] >= 0)
            for (int e = 0; e < 9; ++e) atomicAdd(shydro + 9 * int64_t(hk[s]) + e, hv[s][e]);
}
void lb_hydro_stress(const Box& b, const PartLB* plb, int nPart, double L, double F, double* shydro, hipStream_t s) {
    const int64_t n9 = 9 * int64_t(nPart > 0 ? nPart : 1);
    launch<&k_zero_doubles, 1024, 1>(int((n9 + 255) / 256), 256, s, shydro, n9);
    if (nPart <= 0) return;
    const int64_t threads = int64_t(b.g.X) * b.g.Y * ((b.g.Z + HS_ZRUN - 1) / HS_ZRUN);
    const int nb = int((threads + 255) / 256);
    if (b.prec == LBDEM_PREC_FP64)
        launch<&k_hydro_stress<double>, 1024, 1>(nb, 256, s, b.g, b.type, b.solid, (const double*)b.aux, plb, L, F, shydro);
    else
        launch<&k_hydro_stress<float>, 1024, 1>(nb, 256, s, b.g, b.type, b.solid, (const float*)b.aux, plb, L, F, shydro);
}

void lb_init_fields(const Box& b, const Consts& k, hipStream_t s) {
    const int bs = 256;
    const int nb = int((b.g.N + bs - 1) / bs);
    if (b.prec == LBDEM_PREC_FP64)
        launch<&k_init_fields<double>, 1024, 1>(nb, bs, s, b.g, (double*)b.f, b.type, (double*)b.mass, (double*)b.aux,
                                                (double*)b.visc, k);
    else if (b.prec == LBDEM_PREC_FP32)
        launch<&k_init_fields<float>, 1024, 1>(nb, bs, s, b.g, (float*)b.f, b.type, (float*)b.mass, (float*)b.aux,
                                               (float*)b.visc, k);
    else
        launch<&k_init_fields<__half>, 1024, 1>(nb, bs, s, b.g, (__half*)b.f, b.type, (float*)b.mass, (float*)b.aux,
                                                (float*)b.visc, k);
}

// ------------------------------------------------------------------------------------------
// fused collide + stream
// ------------------------------------------------------------------------------------------
constexpr int FX = FUSED_TILE_X, FY = FUSED_TILE_Y;  // 64x4 threads: 4 row segments, each lane walks FZ cells
                                                     // along z (FZ 1, 2, 4 or 8, chosen per box)
constexpr int HS = 64;                  // per-block particle hash slots
constexpr int PART_STRIDE = PART_STRIDE_;  // per-block partials: obj (7*MAX_OBJ_PART) + walls 18
constexpr int MAX_OBJ_PART = 2;
// block partials of the fused kernel (sPart, doubles): object o's force and torque at 7 o, the
// bounce-back mass term of all object links at 6, wall w's force at 7 MAX_OBJ_PART + 3 w
constexpr int PART_BB = 6, PART_WALL = 7 * MAX_OBJ_PART;

template <typename S>
struct FusedArgs {
    Geom g;
    S* f;
    S* fq[Q];  // population planes (f + q N): uniform bases of the fused kernel's accesses
    uint8_t* type;
    const int32_t* solid;
    typename Types<S>::M* mass;
    typename Types<S>::C* aux;
    typename Types<S>::C* visc;  // node viscosities (the NN variant of the kernel only)
    int32_t* ifList;
    int ifCap;
    const PartLB* plb;
    const ObjLB* olb;
    double* hacc;
    double* blockPart;
    int32_t* blockFlag;
    const uint64_t* bcm;
    DevScalars* sc;
    int nseg, nby, nbz;  // 64-node segments per row (tiles per row with 2-byte storage), row groups, z-slabs
    int gridBlocks;  // the box's grid (a batch launch covers the largest box; the rest return)
    int storeVel;    // every active node's density and velocity into aux (else interface nodes only)
    Consts k;
};

// moving-object link (LB.cpp:1427-1453): bounce-back with the velocity of the object whose sphere
// holds the neighbour (the last one, as the reference scans objects in order), momentum exchange
// into the partials of the object named by the neighbour's solid index oi
template <typename S>
__device__ inline typename Types<S>::C object_value(const FusedArgs<S>& a, int j, int32_t yj, int oi,
                                                   typename Types<S>::C gj, typename Types<S>::C n, double mi,
                                                   double* sPart, const ObjLB* sObj) {
    using C = typename Types<S>::C;
    const int nObj = a.k.nObj;
    int yx, yy, yz;
    cell_xyz(a.g, yj, yx, yy, yz);
    // the impactor node's velocity (LB.cpp:1440): updateObjectBoundaries (LB.cpp:822-837) gives
    // every impactor node a fresh node each step and the velocity of the last impactor whose
    // sphere holds it, or none (zero) when no sphere holds it (a node a restart file typed as
    // impactor away from where the impactor now is)
    int ov = -1;
    for (int oo = 0; oo < nObj; ++oo)
        if (inside_sphere(yx, yy, yz, sObj[oo].c[0], sObj[oo].c[1], sObj[oo].c[2], sObj[oo].r)) ov = oo;
    const double cuo =
        ov < 0 ? 0.0 : sObj[ov].u[0] * cx_rt(j) + sObj[ov].u[1] * cy_rt(j) + sObj[ov].u[2] * cz_rt(j);
    const C BBi = C(6.0) * n * C(w_rt(j)) * C(cuo);
    if (oi >= 0 && oi < nObj && oi < MAX_OBJ_PART) {
        const double fm = 2.0 * double(gj) - double(BBi);
        const double Fb[3] = {fm * cx_rt(j), fm * cy_rt(j), fm * cz_rt(j)};
        const double rr[3] = {yx - sObj[oi].c[0], yy - sObj[oi].c[1], yz - sObj[oi].c[2]};
        const double M[3] = {rr[1] * Fb[2] - rr[2] * Fb[1], rr[2] * Fb[0] - rr[0] * Fb[2], rr[0] * Fb[1] - rr[1] * Fb[0]};
        double* sp = &sPart[7 * oi];
#pragma unroll
        for (int c = 0; c < 3; ++c) { atomicAdd(sp + c, Fb[c]); atomicAdd(sp + 3 + c, M[c]); }
    }
    atomicAdd(&sPart[PART_BB], double(BBi) * mi);
    return gj - BBi;
}

// boundary link of the storing cell (LB.cpp:1390-1453): value reflected into the slot the cell
// itself loads next step; momentum exchange into the block partials
template <typename S, bool DIAG>
__device__ inline typename Types<S>::C boundary_value(const FusedArgs<S>& a, int j, int32_t yj, uint8_t ty,
                                                     typename Types<S>::C gj, typename Types<S>::C n,
                                                     typename Types<S>::C sx, typename Types<S>::C sy,
                                                     typename Types<S>::C sz, typename Types<S>::C usq, double mi,
                                                     double* sPart, const ObjLB* sObj) {
    using C = typename Types<S>::C;
    const Consts& k = a.k;
    if (ty == T_GAS) {
        const C cu = C(cx_rt(j)) * sx + C(cy_rt(j)) * sy + C(cz_rt(j)) * sz;
        return -gj + C(w_rt(j)) * (C(9) * (cu * cu) - C(3) * usq);
    }
    if (ty == T_WALL) {
        if (DIAG && k.nWall > 0 && !k.slip) {
            int wi = a.solid[yj];
            wi = (wi >= 0 && wi < k.nWall) ? wi : 0;
            const double v2 = 2.0 * double(gj);
            if (cx_rt(j)) atomicAdd(&sPart[PART_WALL + 3 * wi + 0], v2 * cx_rt(j));
            if (cy_rt(j)) atomicAdd(&sPart[PART_WALL + 3 * wi + 1], v2 * cy_rt(j));
            if (cz_rt(j)) atomicAdd(&sPart[PART_WALL + 3 * wi + 2], v2 * cz_rt(j));
        }
        return gj;
    }
    if (ty == T_OBJ) return object_value<S>(a, j, yj, a.solid[yj], gj, n, mi, sPart, sObj);
    atomicOr(&a.sc->typeError, 1);
    return gj;
}

// Population addresses in the fused kernel: a uniform part in scalar registers (plane, row and
// the wrapped y/z steps: a wavefront is one x-row segment, so y and z are the same on all its
// lanes) plus a 32-bit per-lane byte offset (x, or x -+ 1 wrapped through periodic x faces) that
// stays fixed along the z-walk. Loads and stores take the scalar-base form: no 64-bit address
// arithmetic per access, no per-lane wrap registers.
struct LaneX {
    uint32_t m, o, p;  // byte offsets of x - 1, x, x + 1 within a row (wrapped, LB.cpp:632-678)
    __device__ LaneX(const Geom& g, int x, uint32_t elt) {
        const int xm = (x == 1 && g.per[0]) ? g.X - 2 : x - 1;
        const int xp = (x == g.X - 2 && g.per[1]) ? 1 : x + 1;
        m = uint32_t(xm) * elt;
        o = uint32_t(x) * elt;
        p = uint32_t(xp) * elt;
    }
    __device__ uint32_t sel(int cx) const { return cx < 0 ? m : (cx > 0 ? p : o); }
};
struct RowYZ {  // uniform across a wavefront
    int32_t row;             // index of node (0, y, z)
    int32_t ym, yp, zm, zp;  // index steps to the wrapped rows y -+ 1 and z -+ 1
    __device__ RowYZ(const Geom& g, int y, int z) {
        const int32_t sy = g.X, sz = g.X * g.Y;
        row = sy * y + sz * z;
        ym = (y == 1 && g.per[2]) ? (g.Y - 3) * sy : -sy;
        yp = (y == g.Y - 2 && g.per[3]) ? -(g.Y - 3) * sy : sy;
        zm = (z == 1 && g.per[4]) ? (g.Z - 3) * sz : -sz;
        zp = (z == g.Z - 2 && g.per[5]) ? -(g.Z - 3) * sz : sz;
    }
    // index of (0, y, z) of the row holding the wrapped neighbour in direction q
    __device__ int32_t nb(int q) const {
        return row + (cyq(q) < 0 ? ym : (cyq(q) > 0 ? yp : 0)) + (czq(q) < 0 ? zm : (czq(q) > 0 ? zp : 0));
    }
};
// element of a plane (uniform base) at a uniform row index plus a per-lane byte offset; a node's
// byte offset within a plane fits 32 bits (19 N < 2^31 nodes, at most 8 bytes each)
template <typename T>
__device__ inline T* at(T* plane, int32_t rowIdx, uint32_t laneBytes) {
    return reinterpret_cast<T*>(reinterpret_cast<char*>(plane) + (uint32_t(rowIdx) * uint32_t(sizeof(T)) + laneBytes));
}

// populations of a node before collision (its post-streaming values). PAR 0: the node's own
// slots; PAR 1: the slots its neighbours wrote (the neighbour in direction opp(q) holds slot
// opp(q)), wrapped through periodic faces
template <typename S, int PAR>
__device__ inline void load_pop(const FusedArgs<S>& a, const RowYZ& r, const LaneX& lx, bool act,
                                typename Types<S>::C* gq) {
    if (!act) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (PAR == 0) gq[q] = StoreT<S>::ld(at(a.fq[q], r.row, lx.o));
        else gq[q] = StoreT<S>::ld(at(a.fq[oppq(q)], r.nb(oppq(q)), lx.sel(cxq(oppq(q)))));
    }
}

// streaming of the post-collision values (LB.cpp:1375-1389): PAR 0 into the node's own slots
// (opposite direction), PAR 1 into the neighbours' slots (wrapped as in load_pop)
template <typename S, int PAR>
__device__ inline void store_pop(const FusedArgs<S>& a, const RowYZ& r, const LaneX& lx, bool act,
                                 const typename Types<S>::C* gq) {
    if (!act) return;
    StoreT<S>::st(at(a.fq[0], r.row, lx.o), gq[0]);
#pragma unroll
    for (int j = 1; j < Q; ++j) {
        if (PAR == 0) StoreT<S>::st(at(a.fq[oppq(j)], r.row, lx.o), gq[j]);
        else StoreT<S>::st(at(a.fq[j], r.nb(j), lx.sel(cxq(j))), gq[j]);
    }
}

// one node's collision (node.cpp:75-207): reconstruct, interior forcing of particle nodes
// (LB.cpp:1919-1962), BGK with Guo forcing. gc holds the post-streaming populations on entry
// and the post-collision ones on return.
template <typename S>
struct CellOut {
    typename Types<S>::C n, sx, sy, sz, usq, hx, hy, hz;
    typename Types<S>::C d[7];  // particle contribution: F(3), M(3), fluid volume
    typename Types<S>::C visc;  // the node's viscosity after the collision (NN variant)
    int key;                    // particle index or -1
};

// apparent viscosity (node.cpp:113-168): the shear-rate tensor from the non-equilibrium part of the
// post-streaming populations, relaxed with the node's previous viscosity nu, its magnitude
// (tinyMatrix::magnitude, myvector.cpp:546-550), the rheology model's viscosity plus the
// Smagorinsky term, clamped as std::max(minVisc, std::min(maxVisc, .)) clamps it (a NaN from 0 / 0
// becomes maxVisc). geq is the equilibrium in offset form, as the collision below uses it.
template <typename C>
__device__ inline C apparent_viscosity(const Consts& k, const C* gc, C n, C sx, C sy, C sz, C usq, C nu) {
    C xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C cu = C(cxq(q)) * sx + C(cyq(q)) * sy + C(czq(q)) * sz;
        const C geq = C(wq(q)) * ((n - C(1)) + n * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
        const C d = gc[q] - geq;
        if (cxq(q)) xx += d;
        if (cyq(q)) yy += d;
        if (czq(q)) zz += d;
        if (cxq(q) && cyq(q)) xy += C(cxq(q) * cyq(q)) * d;
        if (cxq(q) && czq(q)) xz += C(cxq(q) * czq(q)) * d;
        if (cyq(q) && czq(q)) yz += C(cyq(q) * czq(q)) * d;
    }
    const C tau = C(0.5) + C(3) * nu;
    const C sc = C(1.5) / (tau * n);
    xx *= sc; yy *= sc; zz *= sc; xy *= sc; xz *= sc; yz *= sc;
    const C shear = C(2) * sqrt(C(0.5) * (xx * xx + yy * yy + zz * zz + C(2) * (xy * xy + xz * xz + yz * yz)));
    C app;
    if (k.rheo == LBDEM_RHEO_BINGHAM) {
        app = C(k.plasticVisc) + C(k.yieldStress) / shear;
    } else if (k.rheo == LBDEM_RHEO_FRICTIONAL || k.rheo == LBDEM_RHEO_VOELLMY) {
        const C pr = C(0.33333333) * (n - C(1));
        const C pressure = C(0) < pr ? pr : C(0);  // std::max(0.0, .)
        app = k.rheo == LBDEM_RHEO_FRICTIONAL ? C(k.visc) + C(k.frictionCoef) * pressure / shear
                                              : (C(k.frictionCoef) * pressure + usq / C(k.voellmyCoef)) / shear;
    } else {
        app = C(k.visc);
    }
    if (k.turb) app += C(k.turbConst) * shear;
    const C lo = C(k.minVisc), hi = C(k.maxVisc);
    const C m = app < hi ? app : hi;  // std::min(maxVisc, app)
    return lo < m ? m : lo;           // std::max(minVisc, .)
}
template <typename S, bool NN>
__device__ inline void collide_cell(const FusedArgs<S>& a, int32_t sp, uint8_t t, int x, int y, int z, double mi,
                                    typename Types<S>::C* gc, typename Types<S>::C nuIn, CellOut<S>& r) {
    using C = typename Types<S>::C;
    const Consts& k = a.k;
    C sum = 0, jx = 0, jy = 0, jz = 0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        sum += gc[q];
        jx += C(cxq(q)) * gc[q]; jy += C(cyq(q)) * gc[q]; jz += C(czq(q)) * gc[q];
    }
    const C n = C(1) + sum;
    const C inv = C(1) / n;  // one division per node: the quotients below differ by rounding only
    const C ux = jx * inv, uy = jy * inv, uz = jz * inv;
    C hx = 0, hy = 0, hz = 0;
    if (t & T_PMASK) {
        const PartLB P = a.plb[sp];
        const C rx = C(double(x) - P.c[0]), ry = C(double(y) - P.c[1]), rz = C(double(z) - P.c[2]);
        const C wx = C(P.w[0]), wy = C(P.w[1]), wz = C(P.w[2]);
        const C lvx = C(P.v[0]) + (wy * rz - wz * ry);
        const C lvy = C(P.v[1]) + (wz * rx - wx * rz);
        const C lvz = C(P.v[2]) + (wx * ry - wy * rx);
        const C lf = C(mi) * inv;
        const C dfx = lf * (ux - lvx), dfy = lf * (uy - lvy), dfz = lf * (uz - lvz);
        hx = -dfx; hy = -dfy; hz = -dfz;
        r.d[0] = dfx; r.d[1] = dfy; r.d[2] = dfz;
        r.d[3] = ry * dfz - rz * dfy; r.d[4] = rz * dfx - rx * dfz; r.d[5] = rx * dfy - ry * dfx;
        r.d[6] = C(mi);
        r.key = sp;
    }
    const C Fx = C(k.lbF[0]) + hx, Fy = C(k.lbF[1]) + hy, Fz = C(k.lbF[2]) + hz;
    const C sx = ux + C(0.5) * Fx * inv, sy = uy + C(0.5) * Fy * inv, sz = uz + C(0.5) * Fz * inv;
    const C usq = sx * sx + sy * sy + sz * sz;
    const C uF = sx * Fx + sy * Fy + sz * Fz;
    C om = C(k.omega), omf = C(k.omegaf);
    if constexpr (NN) {  // the node's own viscosity, updated first when the model asks for it
        const C nu = k.viscUpdate ? apparent_viscosity<C>(k, gc, n, sx, sy, sz, usq, nuIn) : nuIn;
        om = C(1) / (C(0.5) + C(3) * nu);
        omf = C(1) - C(1) / (C(1) + C(6) * nu);
        r.visc = nu;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const C cu = C(cxq(q)) * sx + C(cyq(q)) * sy + C(czq(q)) * sz;
        const C cF = C(cxq(q)) * Fx + C(cyq(q)) * Fy + C(czq(q)) * Fz;
        const C geq = C(wq(q)) * ((n - C(1)) + n * (C(3) * cu + C(4.5) * cu * cu - C(1.5) * usq));
        gc[q] = gc[q] + om * (geq - gc[q]) + omf * C(wq(q)) * (C(9) * cu * cF + C(3) * (cF - uF));
    }
    r.n = n; r.sx = sx; r.sy = sy; r.sz = sz; r.usq = usq; r.hx = hx; r.hy = hy; r.hz = hz;
}

// one thread's running sums of particle key into the block's table (LDS hash; global atomics
// when the table is full)
template <typename C>
__device__ inline void add_particle(int key, const C* s, int* hkey, double (*hval)[7], double* hacc) {
    int slot = key & (HS - 1);
    int probes = 0;
    for (; probes < HS; ++probes) {
        const int old = atomicCAS(&hkey[slot], -1, key);
        if (old == -1 || old == key) break;
        slot = (slot + 1) & (HS - 1);
    }
    if (probes < HS) {
#pragma unroll
        for (int c = 0; c < 7; ++c) atomicAdd(&hval[slot][c], double(s[c]));
    } else {
        double* hp = hacc + int64_t(key) * 8;
#pragma unroll
        for (int c = 0; c < 7; ++c) atomicAdd(hp + c, double(s[c]));
    }
}

__device__ inline void flush_particles(int tid, int nthreads, const int* hkey, double (*hval)[7], double* hacc) {
    for (int s = tid; s < HS; s += nthreads) {
        const int key = hkey[s];
        if (key >= 0) {
            double* hp = hacc + int64_t(key) * 8;
            for (int c = 0; c < 7; ++c) atomicAdd(hp + c, hval[s][c]);
        }
    }
}

// boundary-link classes of the active node i at (x, y, z), over its wrapped neighbours
__device__ inline uint64_t link_classes(const Geom& g, const uint8_t* type, const int32_t* solid, int32_t i, int x,
                                        int y, int z) {
    const Offs o(g, x, y, z);
    uint64_t m = 0;
    int wprim = -1;
#pragma unroll
    for (int j = 1; j < Q; ++j) {
        const int32_t nb = i + o.d(j);
        const uint8_t tf = type[nb];
        const uint8_t ty = tf & T_BASE;
        uint64_t cls = BC_OTHER;
        if (ty == T_FLUID || ty == T_IF) {
            cls = BC_NONE;
        } else if (ty == T_GAS) {
            cls = BC_GAS;
        } else if (ty == T_WALL) {
            const int w = (tf >> T_WSHIFT) & 7;
            if (wprim < 0) wprim = w;
            cls = w == wprim ? BC_WALL : BC_OTHER;
        } else if (ty == T_OBJ) {
            const int32_t si = solid[nb];
            cls = si == 0 ? BC_OBJ0 : (si == 1 ? BC_OBJ1 : BC_OTHER);
        }
        m |= cls << (BC_BITS * (j - 1));
    }
    if (wprim > 0) m |= uint64_t(wprim) << BC_WALL_SHIFT;
    return m;
}

// Fused collide-stream over the box. Block = FX x FY columns; each thread walks FZ nodes along z.
// Per node the loads of the next node (populations, mass, solid index, boundary-link classes)
// and the type byte of the node after it go out before this node's stores (in either AA parity a
// node loads and stores only its own slot set, so no store of node z aliases a load of node
// z+1). Edge nodes are handled by the same waves as their row neighbours, so every population
// access stays coalesced; their boundary links run in a rolled loop over the precomputed link
// classes, without further loads for gas and wall links.
// Launch bounds of the fused kernel, in blocks (one wavefront per SIMD each) per CU: four for
// 4- and 2-byte storage (128 VGPRs: the free-surface FP32 and FP16 kernels fit without spilling,
// those without a free surface spill up to 4 words; LDS is at most 13 KB per block), two for
// FP64 (its populations alone take 38 VGPRs per node, and two nodes are live).
template <typename S> struct FusedOcc { static constexpr int value = sizeof(S) == 8 ? 2 : 4; };
// Block-to-node mapping. Row order (see k_fused) for 4- and 8-byte storage; measured on the
// 256x320x256 impact box in the impact phase: FP32 0.726 -> 0.631 ms per launch, FP64
// 1.310 -> 1.198 ms against 64 x 4 tiles spread over the XCDs (a probe of the bare AA accesses:
// 5.2 TB/s in either parity in row order, 3.8 TB/s odd / 4.6 TB/s even with the tiles). 2-byte
// storage keeps the tiles (0.521 ms there against 0.656 ms in row order).
template <typename S> struct FusedRowOrder { static constexpr bool value = sizeof(S) >= 4; };
// nodes whose bundle is loaded ahead of the one being collided. Measured as above: none 0.71 ms,
// one 0.69 ms, two 0.78 ms (more loads in flight per CU slow the memory system down).
template <typename S> struct FusedAhead { static constexpr int value = 1; };

// per-node bundle: type byte, populations, mass (particle or edge node), solid index (particle
// node); the boundary-link classes of edge nodes (rare) are loaded where they are used
template <typename S>
struct Bundle {
    typename Types<S>::C g[Q];
    typename Types<S>::M m;
    int32_t s;
    uint8_t t;
    bool act;
};
template <typename S, int PAR>
__device__ inline void load_bundle(const FusedArgs<S>& a, const RowYZ& r, const LaneX& lx, int32_t i, uint8_t t,
                                   Bundle<S>& n) {
    using M = typename Types<S>::M;
    n.t = t;
    n.act = is_active(t);
    load_pop<S, PAR>(a, r, lx, n.act, n.g);
    n.m = (n.act && (t & (T_PMASK | T_EDGE))) ? a.mass[i] : M(0);
    n.s = (n.act && (t & T_PMASK)) ? a.solid[i] : 0;
}

template <typename S, int PAR, bool FSURF, bool DIAG, int FZ, bool NN>
__device__ inline void k_fused(FusedArgs<S> a) {
    using C = typename Types<S>::C;
    using M = typename Types<S>::M;
    constexpr int NT = FX * FY;
    if (int(blockIdx.x) >= a.gridBlocks) return;
    __shared__ int hkey[HS];
    __shared__ double hval[HS][7];
    __shared__ double sPart[PART_STRIDE]; // object / wall partials of the block (FP64 LDS atomics)
    __shared__ int32_t sIf[NT * FZ];      // interface nodes of the block
    __shared__ int sIfCount, sIfBase;
    __shared__ ObjLB sObj[MAX_OBJ_PART];  // the impactors' centres, radii and velocities
    const int tid = threadIdx.x + FX * threadIdx.y;
    if (tid < MAX_OBJ_PART * 8 && tid / 8 < a.k.nObj)
        reinterpret_cast<double*>(sObj)[tid] = reinterpret_cast<const double*>(a.olb)[tid];
    for (int s = tid; s < HS; s += NT) {
        hkey[s] = -1;
        for (int c = 0; c < 7; ++c) hval[s][c] = 0.0;
    }
    if (tid < PART_STRIDE) sPart[tid] = 0.0;
    if (tid == 0) sIfCount = 0;
    __syncthreads();

    static_assert(FX == 64, "a wavefront is one x-row segment: y and z are uniform across it");
    const Geom& g = a.g;
    const int b0 = blockIdx.x;
    int x, y, bz;
    bool col_ok;
    if constexpr (FusedRowOrder<S>::value) {
        // Row order: wavefront threadIdx.y of block b takes row segment s = FY b + threadIdx.y, i.e.
        // segment seg (64 nodes along x) of row y of z-slab bz, s = (bz Y + y) nseg + seg, and walks
        // the slab's FZ nodes in z. The blocks resident at a time then cover a few consecutive
        // z-planes, and the x -+ 1 accesses of a segment's edge lanes meet the neighbouring
        // segment's in the same or the next block while both are in flight.
        const int sid = b0 * FY + int(threadIdx.y);
        const int seg = sid % a.nseg, rest = sid / a.nseg;
        y = rest % g.Y;
        bz = rest / g.Y;
        x = seg * FX + threadIdx.x;
        col_ok = bz < a.nbz && x < g.X;
    } else {
        // 64 x 4 tiles, consecutive tiles on one XCD (tile = (b % 8) per + b / 8)
        const int nblk = a.nseg * a.nby * a.nbz;
        const int per = (nblk + 7) / 8;
        const int tile = (b0 % 8) * per + b0 / 8;
        const int bx = tile % a.nseg, by = (tile / a.nseg) % a.nby;
        bz = tile / (a.nseg * a.nby);
        x = bx * FX + threadIdx.x;
        y = by * FY + threadIdx.y;
        col_ok = tile < nblk && x < g.X && y < g.Y;
    }
    const int yw = __builtin_amdgcn_readfirstlane(y);
    const int32_t N = int32_t(g.N), SZ = g.X * g.Y;
    const int lane = __lane_id();
    const int z0 = bz * FZ;
    auto type_at = [&](int zz) -> uint8_t {
        return (col_ok && zz < FZ && z0 + zz < g.Z) ? a.type[int32_t(cidx(g, x, y, z0 + zz))] : uint8_t(T_WALL);
    };

    // Bundles of the current node and of the AH nodes after it are in registers; per node the
    // thread loads the bundle AH nodes ahead and the type byte one further before it collides and
    // stores the current node. In either AA parity a node loads and stores only its own slot set,
    // so no load of a node ahead reads a slot that a store of the current node writes.
    constexpr int AH = FusedAhead<S>::value;
    Bundle<S> nb[AH + 1];  // nb[0]: current node, nb[k]: k nodes ahead
    int32_t ic = col_ok ? int32_t(cidx(g, x, y, z0)) : 0;
    const LaneX lx(g, x, uint32_t(sizeof(S)));
#pragma unroll
    for (int k = 0; k < AH; ++k) load_bundle<S, PAR>(a, RowYZ(g, yw, z0 + k), lx, ic + k * SZ, type_at(k), nb[k]);
    uint8_t tAhead = type_at(AH);  // type byte of the node AH ahead
    int accKey = -1;  // particle whose contributions acc holds
    C acc[7];
#pragma unroll
    for (int c = 0; c < 7; ++c) acc[c] = 0;
    for (int zz = 0; zz < FZ; ++zz) {
        const int z = z0 + zz;
        const uint8_t tFar = type_at(zz + AH + 1);
        // The wavefront issues its memory operations at raised priority, so that with three
        // wavefronts per SIMD the next node's loads (and the current node's stores) leave ahead of
        // the other wavefronts' arithmetic: measured 1 % faster (0.688 -> 0.679 ms per launch on
        // the 256x320x256 impact box, FP32) than at the default priority throughout.
        __builtin_amdgcn_s_setprio(2);
        load_bundle<S, PAR>(a, RowYZ(g, yw, z + AH), lx, ic + AH * SZ, tAhead, nb[AH]);
        __builtin_amdgcn_s_setprio(0);
        C* gc = nb[0].g;
        const bool actc = nb[0].act;
        const uint8_t tc = nb[0].t;
        const double mc = double(nb[0].m);
        const int32_t sc_ = nb[0].s;

        CellOut<S> r;
        r.key = -1;
#pragma unroll
        for (int c = 0; c < 7; ++c) r.d[c] = 0;
        r.n = 1; r.sx = 0; r.sy = 0; r.sz = 0; r.usq = 0;
        C nu = 0;
        if constexpr (NN) {
            if (actc) nu = a.visc[ic];
        }
        if (actc) collide_cell<S, NN>(a, sc_, tc, x, y, z, mc, gc, nu, r);
        if constexpr (NN) {
            if (actc && a.k.viscUpdate) a.visc[ic] = r.visc;
        }
        __builtin_amdgcn_s_setprio(1);
        store_pop<S, PAR>(a, RowYZ(g, yw, z), lx, actc, gc);
        __builtin_amdgcn_s_setprio(0);

        // boundary links of edge nodes (LB.cpp:1390-1528): the value reflected into the slot
        // this node loads next step
        // The post-collision value of link j is read back from where store_pop just put it
        // (PAR 0: the node's own slot opp(j); PAR 1: slot j of the neighbour), which the same
        // lane wrote, so it needs no staging; the reflected value then goes to the slot the node
        // loads next step.
        const bool edge = actc && (tc & T_EDGE);
        if (__ballot(edge)) {
            const uint64_t bc_ = edge ? a.bcm[ic] : 0;
            uint32_t lm = 0;
            if (edge) {
#pragma unroll
                for (int j = 1; j < Q; ++j)
                    if ((bc_ >> (BC_BITS * (j - 1))) & 7u) lm |= 1u << j;
            }
            const Offs o(g, x, y, z);
            const int wprim = int(bc_ >> BC_WALL_SHIFT) & 7;
            while (lm) {
                const int j = __ffs(lm) - 1;
                lm &= lm - 1u;
                const int cls = int(bc_ >> (BC_BITS * (j - 1))) & 7;
                const int32_t yj = ic + offs_rt(o, j);
                S* const own = a.f + int64_t(opp_rt(j)) * N + ic;  // slot opp(j) of this node
                S* const nbs = a.f + int64_t(j) * N + yj;          // slot j of the neighbour
                const C gj = StoreT<S>::ld(PAR == 0 ? own : nbs);
                C bcv;
                if (cls == BC_GAS) {
                    const C cu = C(cx_rt(j)) * r.sx + C(cy_rt(j)) * r.sy + C(cz_rt(j)) * r.sz;
                    bcv = -gj + C(w_rt(j)) * (C(9) * (cu * cu) - C(3) * r.usq);
                } else if (cls == BC_WALL) {
                    if (DIAG && a.k.nWall > 0 && !a.k.slip) {  // slip walls take no force (LB.cpp:1454-1482)
                        const double v2 = 2.0 * double(gj);
                        double* wp = &sPart[PART_WALL + 3 * wprim];
                        if (cx_rt(j)) atomicAdd(wp + 0, v2 * cx_rt(j));
                        if (cy_rt(j)) atomicAdd(wp + 1, v2 * cy_rt(j));
                        if (cz_rt(j)) atomicAdd(wp + 2, v2 * cz_rt(j));
                    }
                    bcv = gj;
                } else if (cls == BC_OBJ0 || cls == BC_OBJ1) {
                    bcv = object_value<S>(a, j, yj, cls - BC_OBJ0, gj, r.n, mc, sPart, sObj);
                } else {
                    const uint8_t ty = a.type[yj] & T_BASE;
                    if (ty == T_FLUID || ty == T_IF) continue;  // an object node uncovered since marking
                    bcv = boundary_value<S, DIAG>(a, j, yj, ty, gj, r.n, r.sx, r.sy, r.sz, r.usq, mc, sPart, sObj);
                }
                StoreT<S>::st(PAR == 0 ? nbs : own, bcv);
            }
        }
        if (FSURF) {
            const bool isIf = actc && (tc & T_BASE) == T_IF;
            if (actc && !isIf) a.mass[ic] = M(r.n);
            if (isIf || (actc && a.storeVel)) {
                a.aux[ic] = r.n;
                a.aux[N + ic] = r.sx; a.aux[2 * N + ic] = r.sy; a.aux[3 * N + ic] = r.sz;
                a.aux[4 * N + ic] = r.hx; a.aux[5 * N + ic] = r.hy; a.aux[6 * N + ic] = r.hz;
            }
            // interface list: one LDS reservation per wave
            const unsigned long long m = __ballot(isIf);
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                int base = 0;
                if (lane == leader) base = atomicAdd(&sIfCount, __popcll(m));
                base = __shfl(base, leader, 64);
                if (isIf) sIf[base + __popcll(m & ((1ull << lane) - 1ull))] = ic;
            }
        } else if (actc && a.storeVel) {
            a.aux[ic] = r.n;
            a.aux[N + ic] = r.sx; a.aux[2 * N + ic] = r.sy; a.aux[3 * N + ic] = r.sz;
            a.aux[4 * N + ic] = r.hx; a.aux[5 * N + ic] = r.hy; a.aux[6 * N + ic] = r.hz;
        }
        // particle sums: the thread's z-walk usually stays in one sphere for several nodes, so the
        // contributions are summed in registers while the particle index is unchanged and added
        // to the block's table when it changes (and after the walk)
        if (r.key != accKey) {
            if (accKey >= 0) add_particle<C>(accKey, acc, hkey, hval, a.hacc);
            accKey = r.key;
#pragma unroll
            for (int c = 0; c < 7; ++c) acc[c] = r.d[c];
        } else if (r.key >= 0) {
#pragma unroll
            for (int c = 0; c < 7; ++c) acc[c] += r.d[c];
        }
        // advance the column
        ic += SZ;
        tAhead = tFar;
#pragma unroll
        for (int k = 0; k < AH; ++k) nb[k] = nb[k + 1];
    }
    if (accKey >= 0) add_particle<C>(accKey, acc, hkey, hval, a.hacc);
    __syncthreads();
    flush_particles(tid, NT, hkey, hval, a.hacc);
    if (FSURF) {
        // interface list: one global reservation per block
        const int n = sIfCount;
        if (n > 0) {
            if (tid == 0) sIfBase = atomicAdd(&a.sc->nIf, n);
            __syncthreads();
            const int base = sIfBase;
            for (int k = tid; k < n; k += NT) {
                if (base + k < a.ifCap) a.ifList[base + k] = sIf[k];
                else atomicOr(&a.sc->listOverflow, 1);
            }
        }
    }
    // object / wall partials: written only by blocks that have any, flagged for the reduction
    const int any = __syncthreads_or(tid < PART_STRIDE && sPart[tid] != 0.0);
    if (any && tid < PART_STRIDE) a.blockPart[int64_t(tid) * a.gridBlocks + b0] = sPart[tid];
    if (tid == 0) a.blockFlag[b0] = any;
}

// Edge bits (T_EDGE) and boundary-link classes. An active node is an edge node when some wrapped
// neighbour (D3Q19 stencil) is not active. The full pass runs at initialisation, after a particle
// flag rebuild that cleared wall indices, and when the change list overflowed; otherwise only the
// stencil neighbourhoods of the nodes on the change list are re-marked (only transitions between
// active and non-active types, or among non-active types, change edge bits or classes).
//
// Full pass, rows of a multiple of four bytes: one thread per four nodes of a row (one 32-bit
// word of type bytes). Neighbour rows (wrapped through periodic faces) are read as words; the test
// uses three masks of non-active bytes: the node's own row (x - 1, x + 1), the four rows sharing
// one axis with it (x - 1 .. x + 1) and the four diagonal rows (same x).
__device__ inline uint32_t inactive_bytes(uint32_t w) {
    // a byte is active when its type is 0 (fluid) or 3 (interface); types fit in 4 bits, so the
    // per-byte additions below cannot carry into the next byte
    const uint32_t base = w & 0x0F0F0F0Fu;
    const uint32_t nz0 = (base + 0x7F7F7F7Fu) & 0x80808080u;               // type != 0
    const uint32_t nz3 = ((base ^ 0x03030303u) + 0x7F7F7F7Fu) & 0x80808080u;  // type != 3
    uint32_t m = (nz0 & nz3) >> 7;
    m |= m >> 7;
    m |= m >> 14;
    return m & 0xFu;
}
constexpr int MARK_GRID = 1024;  // blocks of k_mark (grid-stride; the full pass is rare)
__device__ __attribute__((always_inline)) inline void mark_full_w(Geom g, uint8_t* type, const int32_t* solid, uint64_t* bcm) {
    const int W = g.X / 4;
    const uint32_t* t32 = reinterpret_cast<const uint32_t*>(type);
    const int64_t total = int64_t(g.Y) * g.Z * W;
    for (int64_t gid = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; gid < total;
         gid += int64_t(gridDim.x) * blockDim.x) {
        const int w = int(gid % W);
        const int64_t row = gid / W;
        const int y = int(row % g.Y), z = int(row / g.Y);
        if (y == 0 || y == g.Y - 1 || z == 0 || z == g.Z - 1) continue;
        const int64_t c0 = row * W + w;
        const uint32_t cw = t32[c0];
        if (inactive_bytes(cw) == 0xFu) continue;  // no active node in this word
        const int ym = (y == 1 && g.per[2]) ? g.Y - 2 : y - 1, yp = (y == g.Y - 2 && g.per[3]) ? 1 : y + 1;
        const int zm = (z == 1 && g.per[4]) ? g.Z - 2 : z - 1, zp = (z == g.Z - 2 && g.per[5]) ? 1 : z + 1;
        auto rowWord = [&](int yy, int zz) { return (int64_t(zz) * g.Y + yy) * W; };
        // 6-bit window of a row: bit 0 = byte x-1, bits 1-4 = the word, bit 5 = byte x+4; bytes of
        // the boundary layer of a periodic x face are replaced by their wrapped partners
        auto window = [&](int64_t rw) {
            uint32_t mid = inactive_bytes(t32[rw + w]);
            if (w == 0 && g.per[0]) {
                const uint32_t b = (t32[rw + W - 1] >> 16) & 0xFFu;  // x = X - 2
                mid = (mid & ~1u) | (is_active(uint8_t(b)) ? 0u : 1u);
            }
            if (w == W - 1 && g.per[1]) {
                const uint32_t b = (t32[rw] >> 8) & 0xFFu;  // x = 1
                mid = (mid & ~8u) | (is_active(uint8_t(b)) ? 0u : 8u);
            }
            const uint32_t lo = w > 0 ? uint32_t(!is_active(uint8_t(t32[rw + w - 1] >> 24))) : 1u;
            const uint32_t hi = w < W - 1 ? uint32_t(!is_active(uint8_t(t32[rw + w + 1]))) : 1u;
            return lo | (mid << 1) | (hi << 5);
        };
        const uint32_t own = window(rowWord(y, z));
        const uint32_t side = window(rowWord(ym, z)) | window(rowWord(yp, z)) | window(rowWord(y, zm)) |
                              window(rowWord(y, zp));
        const uint32_t diag = inactive_bytes(t32[rowWord(ym, zm) + w]) | inactive_bytes(t32[rowWord(ym, zp) + w]) |
                              inactive_bytes(t32[rowWord(yp, zm) + w]) | inactive_bytes(t32[rowWord(yp, zp) + w]);
        uint32_t nw = cw;
        // rolled: one inlined copy of the link-class code keeps the kernel free of calls and scratch
#pragma unroll 1
        for (int k = 0; k < 4; ++k) {
            const uint8_t t = uint8_t(cw >> (8 * k));
            if (!is_active(t)) continue;
            const bool edge = ((own >> k) & 5u) || ((side >> k) & 7u) || ((diag >> k) & 1u);
            const uint8_t nt = edge ? uint8_t(t | T_EDGE) : uint8_t(t & ~T_EDGE);
            nw = (nw & ~(0xFFu << (8 * k))) | (uint32_t(nt) << (8 * k));
            if (edge) bcm[c0 * 4 + k] = link_classes(g, type, solid, int32_t(c0 * 4 + k), 4 * w + k, y, z);
        }
        // other threads read this word concurrently; they only test the type bits, which do not change
        if (nw != cw) reinterpret_cast<uint32_t*>(type)[c0] = nw;
    }
}
// one active node: edge bit and classes from its wrapped neighbours
__device__ inline void mark_node(const Geom& g, uint8_t* type, const int32_t* solid, uint64_t* bcm, int32_t i) {
    const uint8_t t = type[i];
    if (!is_active(t)) return;
    int x, y, z;
    cell_xyz(g, i, x, y, z);
    const uint64_t m = link_classes(g, type, solid, i, x, y, z);
    const bool edge = m != 0;
    if (edge) bcm[i] = m;
    const uint8_t nt = edge ? uint8_t(t | T_EDGE) : uint8_t(t & ~T_EDGE);
    if (nt != t) type[i] = nt;
}
// full pass for rows whose length is not a multiple of four: one thread per node
__device__ inline void mark_full_b(Geom g, uint8_t* type, const int32_t* solid, uint64_t* bcm) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x)
        mark_node(g, type, solid, bcm, int32_t(i));
}
// One launch: the full pass when the full-pass flag is set (initialisation, a flag rebuild that
// clears wall indices, change-list overflow), else the incremental pass: the stencil
// neighbourhood (wrapped) of every node on the change list. A node reached from several changed
// nodes is marked several times with the same result.
__device__ inline void k_mark(Geom g, uint8_t* type, const int32_t* solid, uint64_t* bcm, const int32_t* changed, int cap,
                              const DevScalars* sc) {
    if (sc->fullMark) {
        if (g.X % 4 == 0) mark_full_w(g, type, solid, bcm);
        else mark_full_b(g, type, solid, bcm);
        return;
    }
    const int n = min(sc->nChanged, cap);
    for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < int64_t(n) * Q;
         t += int64_t(gridDim.x) * blockDim.x) {
        const int32_t c = changed[t / Q];
        const int j = int(t % Q);
        int x, y, z;
        cell_xyz(g, c, x, y, z);
        const Offs o(g, x, y, z);
        const int64_t p = int64_t(c) + offs_rt(o, j);
        if (p >= 0 && p < g.N) mark_node(g, type, solid, bcm, int32_t(p));  // (changed boundary-layer nodes)
    }
}

void lb_mark_edges(const Box& b, DevScalars* sc, hipStream_t s) {
    launch_gs<&k_mark, 256, 1>(MARK_GRID, 256, s, b.g, b.type, b.solid, b.bcm, b.changed, b.changedCap, sc);
    // the list and the full-pass flag are reset by the next kernel (k_begin_lb)
}

int fused_grid_blocks(const Geom& g, int fz, int prec) {
    const int nseg = (g.X + FX - 1) / FX, nby = (g.Y + FY - 1) / FY, nbz = (g.Z + fz - 1) / fz;
    if (prec != LBDEM_PREC_FP16) return (nseg * g.Y * nbz + FY - 1) / FY;  // row order
    return ((nseg * nby * nbz + 7) / 8) * 8;                              // XCD-spread tiles
}

// Slip walls (LB.cpp:1454-1482, slip static walls): a diagonal link j (j > 6) of an active node A
// into a wall, with exactly one of its two check nodes A + c(slip1Check[j]), A + c(slip2Check[j])
// active, receives S1 times that active check node B's post-collision population k = slip1[j]
// (or slip2[j]) plus S2 = 1 - S1 times A's own post-collision population j; every other wall link
// bounces back, as the fused kernel has already written. Runs right after the fused kernel of
// parity PAR, over the nodes next to a wall: it reads post-collision values where that kernel
// stored them (PAR 0: slot opp(k) of the node itself; PAR 1: slot k of the node's neighbour in
// direction k, here always the inactive check node or the wall) and overwrites the slot A loads
// next step (PAR 0: slot j of the wall node; PAR 1: A's own slot opp(j)), none of which is read here.
__host__ __device__ constexpr int slip1Check(int j) {
    constexpr int c[Q] = {0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 4, 3, 4, 5, 6, 1, 2, 6, 5};
    return c[j];
}
__host__ __device__ constexpr int slip1Dir(int j) {
    constexpr int c[Q] = {0, 0, 0, 0, 0, 0, 0, 9, 10, 8, 7, 13, 14, 12, 11, 18, 17, 15, 16};
    return c[j];
}
__host__ __device__ constexpr int slip2Check(int j) {
    constexpr int c[Q] = {0, 0, 0, 0, 0, 0, 0, 3, 4, 2, 1, 5, 6, 4, 3, 5, 6, 1, 2};
    return c[j];
}
__host__ __device__ constexpr int slip2Dir(int j) {
    constexpr int c[Q] = {0, 0, 0, 0, 0, 0, 0, 10, 9, 7, 8, 14, 13, 11, 12, 17, 18, 16, 15};
    return c[j];
}
template <typename S, int PAR>
__device__ inline void k_slip(Geom g, S* f, const uint8_t* type, const int32_t* list, int n, double s1) {
    using C = typename StoreT<S>::C;
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    const int32_t a = list[i];
    if (!is_active(type[a])) return;
    int x, y, z;
    cell_xyz(g, a, x, y, z);
    const Offs oa(g, x, y, z);
    const int64_t N = g.N;
#pragma unroll
    for (int j = 7; j < Q; ++j) {
        const int32_t w = a + oa.d(j);
        if ((type[w] & T_BASE) != T_WALL) continue;
        const bool act1 = is_active(type[a + oa.d(slip1Check(j))]);
        const bool act2 = is_active(type[a + oa.d(slip2Check(j))]);
        if (act1 == act2) continue;  // bounce-back, written by the fused kernel
        const int32_t b = a + oa.d(act1 ? slip1Check(j) : slip2Check(j));
        const int kk = act1 ? slip1Dir(j) : slip2Dir(j);
        C fb, fa;
        if (PAR == 0) {
            fb = StoreT<S>::ld(f + int64_t(oppq(kk)) * N + b);
            fa = StoreT<S>::ld(f + int64_t(oppq(j)) * N + a);
        } else {
            int bx, by, bz;
            cell_xyz(g, b, bx, by, bz);
            const Offs ob(g, bx, by, bz);
            fb = StoreT<S>::ld(f + int64_t(kk) * N + b + ob.d(kk));
            fa = StoreT<S>::ld(f + int64_t(j) * N + w);
        }
        // offset storage: S1 (fb + w) + S2 (fa + w) - w = S1 fb + S2 fa (the weights of j and kk agree)
        const C v = C(s1) * fb + (C(1) - C(s1)) * fa;
        if (PAR == 0) StoreT<S>::st(f + int64_t(j) * N + w, v);
        else StoreT<S>::st(f + int64_t(oppq(j)) * N + a, v);
    }
}

template <typename S>
static void launch_slip(const Box& b, double s1, hipStream_t s) {
    const int bs = 256, nb = (b.nSlip + bs - 1) / bs;
    if (b.parity == 0) launch<&k_slip<S, 0>, 256, 1>(nb, bs, s, b.g, (S*)b.f, b.type, b.slipList, b.nSlip, s1);
    else launch<&k_slip<S, 1>, 256, 1>(nb, bs, s, b.g, (S*)b.f, b.type, b.slipList, b.nSlip, s1);
}

template <typename S>
static void launch_fused(Box& b, const Consts& k, const PartLB* plb, const ObjLB* olb, DevScalars* sc, bool diag,
                         bool storeVel, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    FusedArgs<S> a;
    a.g = b.g; a.f = (S*)b.f; a.type = b.type; a.solid = b.solid;
    for (int q = 0; q < Q; ++q) a.fq[q] = a.f + int64_t(q) * b.g.N;
    a.mass = (typename Types<S>::M*)b.mass; a.aux = (typename Types<S>::C*)b.aux;
    a.visc = (typename Types<S>::C*)b.visc;
    a.ifList = b.ifList; a.ifCap = b.ifCap;
    a.plb = plb; a.olb = olb; a.hacc = b.hacc; a.blockPart = b.blockPart; a.blockFlag = b.blockFlag; a.bcm = b.bcm; a.sc = sc;
    a.k = k;
    const int fz = b.fz;
    a.nseg = (b.g.X + FX - 1) / FX; a.nby = (b.g.Y + FY - 1) / FY; a.nbz = (b.g.Z + fz - 1) / fz;
    const int grid = fused_grid_blocks(b.g, fz, b.prec);
    a.gridBlocks = grid;
    a.storeVel = storeVel ? 1 : 0;
    const bool fs = k.freeSurface != 0;
    if (ev0) (void)hipEventRecord(ev0, s);
    (void)diag;  // the wall diagnostics are always accumulated
    const bool nn = b.visc != nullptr;
#define LBDEM_FUSED(P, F, Z, V) \
    launch<&k_fused<S, P, F, true, Z, V>, FX * FY, FusedOcc<S>::value>(dim3(grid), dim3(FX, FY, 1), s, a)
#define LBDEM_FUSED_Z(P, F, V)                 \
    switch (fz) {                             \
        case 1: LBDEM_FUSED(P, F, 1, V); break; \
        case 2: LBDEM_FUSED(P, F, 2, V); break; \
        case 4: LBDEM_FUSED(P, F, 4, V); break; \
        default: LBDEM_FUSED(P, F, 8, V); break; \
    }
#define LBDEM_FUSED_FS(P, V) \
    if (fs) { LBDEM_FUSED_Z(P, true, V); } else { LBDEM_FUSED_Z(P, false, V); }
    if (nn) {
        if (b.parity == 0) { LBDEM_FUSED_FS(0, true) } else { LBDEM_FUSED_FS(1, true) }
    } else {
        if (b.parity == 0) { LBDEM_FUSED_FS(0, false) } else { LBDEM_FUSED_FS(1, false) }
    }
#undef LBDEM_FUSED_FS
#undef LBDEM_FUSED_Z
#undef LBDEM_FUSED
    if (ev1) (void)hipEventRecord(ev1, s);
    if (b.slipList && b.nSlip > 0) launch_slip<S>(b, k.slipCoef, s);
    b.parity ^= 1;
}

// end of findNewActive's scan (LB.cpp:1980-2029): the lowest node still inside its particle
// becomes the next step's starting point (-1: none)
__device__ inline void fna_finish(DevScalars* sc) {
    if (!sc->needScan) return;
    sc->kminP = sc->kstar == INT_MAX ? -1 : sc->kstar;
    sc->needScan = 0;
}

// after the edge marking, before the fused kernel: the findNewActive scan's result is taken, the
// change list is consumed, the interface list starts empty (the particle sums were zeroed by the previous step's k_hydro_finalize, or
// at initialisation)
__device__ inline void k_begin_lb(DevScalars* sc) {
    fna_finish(sc);
    sc->fullMark = 0;
    sc->nChanged = 0;
    sc->nIf = 0;
}

void lb_fused_step(Box& b, const Consts& k, const PartLB* plb, const ObjLB* olb, DevScalars* sc, bool diag,
                   bool storeVel, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    launch<&k_begin_lb, 1024, 1>(1, 1, s, sc);
    if (b.prec == LBDEM_PREC_FP64) launch_fused<double>(b, k, plb, olb, sc, diag, storeVel, s, ev0, ev1);
    else if (b.prec == LBDEM_PREC_FP32) launch_fused<float>(b, k, plb, olb, sc, diag, storeVel, s, ev0, ev1);
    else launch_fused<__half>(b, k, plb, olb, sc, diag, storeVel, s, ev0, ev1);
}

// deterministic sum of the flagged per-block partials, unit conversion of all hydrodynamic results
// (LB.cpp:1966-1977, 1536-1543)
constexpr int HF_THREADS = 1024;
__device__ inline void k_hydro_finalize(int nblk, const double* part, const int32_t* flag, Consts k, double* hacc, double* FHydro,
                                 double* MHydro, double* fluidVol, DevScalars* sc, double* objFH, double* objMH,
                                 double* wallFH, int diag, int* count2) {
    __shared__ double red[HF_THREADS / 64];
    const int tid = threadIdx.x;
    if (blockIdx.x == 0 && tid == 0) {  // counters of the free-surface pass that follows (LB.cpp:422-430)
        sc->surplus = 0.0;
        sc->nFilled = sc->nEmptied = sc->nNewIf = 0;
        sc->nIfFinal = 0;
        *count2 = 0;
    }
    if (blockIdx.x < PART_STRIDE) {
        // column c of the partial table (one block per column), summed in a fixed order: strided
        // per-thread sums, a wavefront butterfly, then the waves in order
        const int c = blockIdx.x;
        if (c >= 7 * MAX_OBJ_PART && !diag) return;
        if (c < 7 * MAX_OBJ_PART && c / 7 >= k.nObj && c != 6) return;            // no such object
        if (c >= 7 * MAX_OBJ_PART && (c - 7 * MAX_OBJ_PART) / 3 >= k.nWall) return;  // no such wall
        // four flag loads (and then four value loads) in flight per thread; the additions keep the
        // order of one strided pass (a skipped block adds +0.0, which leaves the sum unchanged)
        double acc = 0.0;
        const int bs = int(blockDim.x);
        const double* col = part + int64_t(c) * nblk;
        int b = tid;
        for (; b + 3 * bs < nblk; b += 4 * bs) {
            const int32_t f0 = flag[b], f1 = flag[b + bs], f2 = flag[b + 2 * bs], f3 = flag[b + 3 * bs];
            const double v0 = f0 ? col[b] : 0.0, v1 = f1 ? col[b + bs] : 0.0;
            const double v2 = f2 ? col[b + 2 * bs] : 0.0, v3 = f3 ? col[b + 3 * bs] : 0.0;
            acc += v0; acc += v1; acc += v2; acc += v3;
        }
        for (; b < nblk; b += bs)
            if (flag[b]) acc += col[b];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if ((tid & 63) == 0) red[tid / 64] = acc;
        __syncthreads();
        if (tid == 0) {
            double v = 0.0;
            for (int w = 0; w < HF_THREADS / 64; ++w) v += red[w];
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
            double* h = hacc + int64_t(p) * 8;
            for (int c = 0; c < 3; ++c) {
                FHydro[3 * p + c] = h[c] * k.unitForce;
                MHydro[3 * p + c] = h[3 + c] * k.unitTorque;
            }
            fluidVol[p] = h[6] * k.unitVolume;
            for (int c = 0; c < 8; ++c) h[c] = 0.0;  // zero for the next LB step's sums
        }
    }
}

void lb_hydro_finalize(const Box& b, const Consts& k, double* FHydro, double* MHydro, double* fluidVol,
                       DevScalars* sc, double* objFH, double* objMH, double* wallFH, bool diag, hipStream_t s) {
    const int nblk = fused_grid_blocks(b.g, b.fz, b.prec);
    const int grid = PART_STRIDE + (k.nPart + HF_THREADS - 1) / HF_THREADS;
    int* count2 = (int*)(b.ifList2 + b.ifCap);  // free-surface counter stored after the list
    launch<&k_hydro_finalize, HF_THREADS, 1>(grid, HF_THREADS, s, nblk, b.blockPart, b.blockFlag, k, b.hacc, FHydro, MHydro, fluidVol, sc, objFH, objMH, wallFH, diag ? 1 : 0, count2);
}

// ------------------------------------------------------------------------------------------
// free surface (LB.cpp:1548-1852)
// ------------------------------------------------------------------------------------------
constexpr int LIST_GRID = 256;  // blocks of the free-surface list passes (few items: launch ramp dominates)

template <typename S, int L>
__device__ inline void k_fs_mass(Geom g, const S* f, const uint8_t* type, const typename Types<S>::M* mass,
                          const typename Types<S>::C* aux, const int32_t* ifList, typename Types<S>::M* newMass,
                          DevScalars* sc, int32_t* filledFlag, int slipDyn, double s1) {
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
        // every neighbour's type, populations and mass are loaded up front (all addresses are
        // nodes of the box), so the loads overlap; the sum then runs in direction order
        uint8_t ty[Q];
        C a1[Q], a2[Q], mj[Q];
#pragma unroll
        for (int j = 1; j < Q; ++j) {
            const int64_t yj = i + o.d(j);
            ty[j] = type[yj] & T_BASE;
            a1[j] = StoreT<S>::ld(f + int64_t(j) * g.N + yj);
            a2[j] = StoreT<S>::ld(f + int64_t(oppq(j)) * g.N + i);
            mj[j] = C(mass[yj]);
        }
        C dm = 0;
#pragma unroll
        for (int j = 1; j < Q; ++j) {
            const bool slipLink = slipDyn && ty[j] == T_WALL;
            if (ty[j] != T_IF && ty[j] != T_FLUID && ty[j] != T_OBJ && !slipLink) continue;
            const C ms = L == 1 ? (a1[j] - a2[j]) : (a2[j] - a1[j]);  // incoming(opp j) - outgoing(j), offsets
            if (ty[j] == T_IF) dm += C(0.5) * (C(double(mj[j]) + addE) + mi) * ms;
            else if (ty[j] == T_FLUID) dm += ms;
            else if (slipLink) {  // moving slip wall (LB.cpp:1595-1625): (1 - S) of the mass on a slip link
                bool slipped = false;
                if (j > 6) {
                    const uint8_t t1 = ty[slip1Check(j)], t2 = ty[slip2Check(j)];
                    slipped = (t1 == T_FLUID || t1 == T_IF) != (t2 == T_FLUID || t2 == T_IF);
                }
                dm += (slipped ? C(1.0 - s1) * mi : mi) * ms;
            } else {
                dm += mi * ms;
            }
        }
        newMass[k] = M(mi + dm);
    }
}

// Position in a list for each lane of the wavefront that calls this now (any subset of its
// lanes, in divergent code too): one atomic reservation per wavefront, lanes in lane order.
__device__ inline int wave_reserve(int* count) {
    const unsigned long long m = __ballot(1);
    const int lane = __lane_id();
    const int leader = __ffsll(m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader, 64);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

// a node changed between active and non-active (or among non-active types): its neighbourhood is
// re-marked before the next LB step; an overflowing list asks for a full re-mark instead
__device__ inline void push_changed(int32_t* changed, int cap, DevScalars* sc, int32_t i) {
    const int pos = wave_reserve(&sc->nChanged);
    if (pos < cap) changed[pos] = i;
    else atomicOr(&sc->fullMark, 1);
}

__device__ inline void list_push(int32_t* list, int* count, int cap, int32_t v, int* overflow) {
    const int pos = wave_reserve(count);
    if (pos < cap) list[pos] = v;
    else atomicOr(overflow, 1);
}

// Block-wide forms of the two above for the passes over the whole interface list, where one
// atomic per wavefront on one counter serialises (every thread of the block must call them, the
// same number of times; sTmp: blockDim.x / 64 + 1 ints or doubles of LDS, reused by the next call):
// positions in a list for the threads with `want`, one reservation per block, block-major then
// wave then lane order; and a sum into one global total, one atomic per block.
__device__ inline int block_reserve(int* count, bool want, int* sTmp) {
    const unsigned long long m = __ballot(want);
    const int lane = __lane_id(), wave = threadIdx.x / 64, nw = blockDim.x / 64;
    if (lane == 0) sTmp[wave] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; ++w) {
            const int c = sTmp[w];
            sTmp[w] = acc;
            acc += c;
        }
        sTmp[nw] = acc ? atomicAdd(count, acc) : 0;
    }
    __syncthreads();
    const int r = sTmp[nw] + sTmp[wave] + __popcll(m & ((1ull << lane) - 1ull));
    __syncthreads();
    return r;
}
template <typename T>
__device__ inline void block_add(T* dst, T v, T* sTmp) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int wave = threadIdx.x / 64, nw = blockDim.x / 64;
    if (__lane_id() == 0) sTmp[wave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        T t = 0;
        for (int w = 0; w < nw; ++w) t += sTmp[w];
        if (t != T(0)) atomicAdd(dst, t);
    }
    __syncthreads();
}

// per-thread partial sums into one global total: summed over the wavefront first (every lane of
// the wavefront must call this), one atomic per wavefront that has a nonzero sum
__device__ inline void wave_add(double* dst, double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (__lane_id() == 0 && v != 0.0) atomicAdd(dst, v);
}
__device__ inline void wave_add(int* dst, int v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (__lane_id() == 0 && v != 0) atomicAdd(dst, v);
}

// assign new mass, classify filled / emptied (findInterfaceMutants, LB.cpp:1669-1698)
template <typename S>
__device__ inline void k_fs_mutants(uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux, int64_t N,
                             const int32_t* ifList, const typename Types<S>::M* newMass, int32_t* filled,
                             int32_t* emptied, int cap, DevScalars* sc, int32_t* marker, int32_t* changed,
                             int changedCap) {
    const int nIf = sc->nIf;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nIf; k += gridDim.x * blockDim.x) {
        const int64_t i = ifList[k];
        const auto m = newMass[k];
        mass[i] = m;
        const auto n = aux[i];
        if (m > n) {
            type[i] = (type[i] & (T_PMASK | T_EDGE)) | T_FLUID;  // active to active: edge bit stays
            marker[i] = -3;
            list_push(filled, &sc->nFilled, cap, int32_t(i), &sc->listOverflow);
        } else if (m < 0) {
            type[i] = (type[i] & T_PMASK) | T_GAS;
            list_push(emptied, &sc->nEmptied, cap, int32_t(i), &sc->listOverflow);
            push_changed(changed, changedCap, sc, int32_t(i));
        }
    }
}

// gas neighbours of filled nodes: donor = highest-index filled neighbour (LB.cpp:1704-1728)
__device__ inline void k_fs_donors(Geom g, const uint8_t* type, const int32_t* filled, int32_t* donor, int32_t* newIf,
                            int cap, DevScalars* sc) {
    const int nF = sc->nFilled;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nF; k += gridDim.x * blockDim.x) {
        const int64_t i = filled[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        uint32_t gasMask = 0;  // neighbour types loaded up front so the loads overlap
#pragma unroll
        for (int j = 1; j < Q; ++j) gasMask |= uint32_t((type[i + o.d(j)] & T_BASE) == T_GAS) << j;
        for (uint32_t m = gasMask; m; m &= m - 1u) {  // gas neighbours in direction order
            const int j = __ffs(m) - 1;
            const int64_t yj = i + o.d(j);
            const int old = atomicMax(&donor[yj], int32_t(i));
            if (old < 0) list_push(newIf, &sc->nNewIf, cap, int32_t(yj), &sc->listOverflow);
        }
    }
}

// initialise the new interface nodes from their donor (node::initialize with the donor's
// u, visc and hydroForce+lbF, mass 0.01*rho0)
template <typename S, int L>
__device__ inline void k_fs_new_from_donor(Geom g, S* f, uint8_t* type, typename Types<S>::M* mass,
                                    typename Types<S>::C* aux, typename Types<S>::C* visc, const int32_t* newIf,
                                    const int32_t* donor, DevScalars* sc, Consts k, int32_t* changed, int changedCap) {
    using C = typename Types<S>::C;
    const int nN = sc->nNewIf;
    for (int kk = blockIdx.x * blockDim.x + threadIdx.x; kk < nN; kk += gridDim.x * blockDim.x) {
        const int64_t gi = newIf[kk];
        const int64_t d = donor[gi];
        const int64_t N = g.N;
        type[gi] = (type[gi] & T_PMASK) | T_IF;
        push_changed(changed, changedCap, sc, int32_t(gi));
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
        if (visc) visc[gi] = visc[d];  // the donor's viscosity (LB.cpp:1719)
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&sc->surplus, -0.01 * k.rho0 * double(nN));
}

// fluid neighbours of emptied nodes become interface with 99% of their density (LB.cpp:1730-1748)
template <typename S>
__device__ inline void k_fs_from_emptied(Geom g, uint8_t* type, typename Types<S>::M* mass, typename Types<S>::C* aux,
                                  const int32_t* emptied, int32_t* marker, int32_t* newIf, int cap, DevScalars* sc) {
    using C = typename Types<S>::C;
    using M = typename Types<S>::M;
    const int nE = sc->nEmptied;
    double gained = 0.0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nE; k += gridDim.x * blockDim.x) {
        const int64_t i = emptied[k];
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        // neighbour types loaded up front so the loads overlap; the marker CAS below decides which
        // thread converts a node, so a type read early changes nothing
        uint32_t fluidMask = 0;
#pragma unroll
        for (int j = 0; j < Q; ++j) fluidMask |= uint32_t((type[i + o.d(j)] & T_BASE) == T_FLUID) << j;
        for (uint32_t m = fluidMask; m; m &= m - 1u) {  // fluid neighbours in direction order
            const int j = __ffs(m) - 1;
            const int64_t l = i + o.d(j);
            double nl;
            int old = atomicCAS(&marker[l], -1, -2);
            if (old == -1) {
                nl = double(mass[l]);               // ordinary fluid node: mass holds its density
            } else if (old == -3 && atomicCAS(&marker[l], -3, -4) == -3) {
                nl = double(aux[l]);                // filled this step: density kept in aux
            } else {
                continue;
            }
            type[l] = (type[l] & (T_PMASK | T_EDGE)) | T_IF;  // active to active: edge bit stays
            mass[l] = M(0.99 * nl);
            aux[l] = C(nl);
            gained += 0.01 * nl;
            list_push(newIf, &sc->nNewIf, cap, int32_t(l), &sc->listOverflow);
        }
    }
    wave_add(&sc->surplus, gained);
}

// updateMutants (LB.cpp:1751-1790)
template <typename S>
__device__ inline void k_fs_update_mutants(uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux,
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
    wave_add(&sc->surplus, acc);
}

// removeIsolated part 1: interface nodes without gas neighbours become fluid (LB.cpp:1796-1816).
// The interface list after the conversions (the old interface nodes left untouched, then the new
// ones) is visited here and compacted into list2 for the two passes that follow. A node's filter
// depends only on its own type and marker, which no other thread changes in this pass, and a
// conversion to fluid changes no node's gas test, so the pass equals the sequential loop.
template <typename S>
__device__ inline void k_fs_isolated_a(Geom g, uint8_t* type, typename Types<S>::M* mass, const typename Types<S>::C* aux,
                                const int32_t* ifList, const int32_t* newIf, const int32_t* marker, int32_t* list2,
                                int cap, int* count2, DevScalars* sc) {
    const int nIf = sc->nIf, nN = sc->nNewIf;
    __shared__ int sPos[8];
    __shared__ double sSum[8];
    double acc = 0.0;
    // the loop runs the same number of times on every thread of a block (block_reserve)
    for (int k0 = blockIdx.x * blockDim.x; k0 < nIf + nN; k0 += gridDim.x * blockDim.x) {
        const int k = k0 + int(threadIdx.x);
        int64_t i = -1;
        if (k < nIf) {
            i = ifList[k];
            if ((type[i] & T_BASE) != T_IF || marker[i] != -1) i = -1;
        } else if (k < nIf + nN) {
            i = newIf[k - nIf];
        }
        bool gasNb = false;
        if (i >= 0) {
            int x, y, z;
            cell_xyz(g, i, x, y, z);
            const Offs o(g, x, y, z);
#pragma unroll
            for (int j = 1; j < Q; ++j) gasNb |= (type[i + o.d(j)] & T_BASE) == T_GAS;  // the 18 loads overlap
        }
        const int pos = block_reserve(count2, i >= 0, sPos);
        if (i >= 0) {
            if (pos < cap) list2[pos] = int32_t(i);
            else atomicOr(&sc->listOverflow, 1);
            if (!gasNb) {
                acc += double(mass[i]) - double(aux[i]);
                mass[i] = typename Types<S>::M(aux[i]);
                type[i] = (type[i] & (T_PMASK | T_EDGE)) | T_FLUID;  // active to active: edge bit stays
            }
        }
    }
    block_add(&sc->surplus, acc, sSum);
}

// removeIsolated part 2: interface nodes without fluid neighbours become gas (LB.cpp:1821-1841)
template <typename S>
__device__ inline void k_fs_isolated_b(Geom g, uint8_t* type, const typename Types<S>::M* mass, const int32_t* list2,
                                const int* count2, DevScalars* sc, int32_t* changed, int changedCap) {
    const int n2 = *count2;
    __shared__ double sSum[8];
    __shared__ int sCnt[8];
    double acc = 0.0;
    int keep = 0;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += gridDim.x * blockDim.x) {
        const int64_t i = list2[k];
        if ((type[i] & T_BASE) != T_IF) continue;
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const Offs o(g, x, y, z);
        bool fluidNb = false;  // the 18 type loads overlap (no early exit)
#pragma unroll
        for (int j = 1; j < Q; ++j) fluidNb |= (type[i + o.d(j)] & T_BASE) == T_FLUID;
        if (!fluidNb) {
            acc += double(mass[i]);
            type[i] = (type[i] & T_PMASK) | T_GAS;
            push_changed(changed, changedCap, sc, int32_t(i));
        } else {
            keep++;
        }
    }
    block_add(&sc->surplus, acc, sSum);
    block_add(&sc->nIfFinal, keep, sCnt);
}

// redistributeMass (LB.cpp:1844-1852) and marker reset
template <typename S>
__device__ inline void k_fs_redistribute(const uint8_t* type, typename Types<S>::M* mass, const int32_t* list2,
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


template <typename S>
static void fs_impl(Box& b, const Consts& k, DevScalars* sc, hipStream_t s) {
    using M = typename Types<S>::M;
    using C = typename Types<S>::C;
    S* f = (S*)b.f;
    M* mass = (M*)b.mass;
    C* aux = (C*)b.aux;
    int* count2 = (int*)(b.ifList2 + b.ifCap);  // counter stored after the list
    const int grid = LIST_GRID, bs = 256;
    if (b.parity == 1)
        launch_gs<&k_fs_mass<S, 1>, 256, 1>(grid, bs, s, b.g, f, b.type, mass, aux, b.ifList, (M*)b.ifNewMass, sc, b.donor,
                                            k.slip == 2 ? 1 : 0, k.slipCoef);
    else
        launch_gs<&k_fs_mass<S, 0>, 256, 1>(grid, bs, s, b.g, f, b.type, mass, aux, b.ifList, (M*)b.ifNewMass, sc, b.donor,
                                            k.slip == 2 ? 1 : 0, k.slipCoef);
    launch_gs<&k_fs_mutants<S>, 256, 1>(grid, bs, s, b.type, mass, aux, b.g.N, b.ifList, (M*)b.ifNewMass, b.filled, b.emptied, b.ifCap, sc, b.donor, b.changed, b.changedCap);
    launch_gs<&k_fs_donors, 256, 1>(grid, bs, s, b.g, b.type, b.filled, b.donor, b.newIf, b.ifCap, sc);
    if (b.parity == 1)
        launch_gs<&k_fs_new_from_donor<S, 1>, 256, 1>(grid, bs, s, b.g, f, b.type, mass, aux, (C*)b.visc, b.newIf, b.donor, sc, k,
                                                      b.changed, b.changedCap);
    else
        launch_gs<&k_fs_new_from_donor<S, 0>, 256, 1>(grid, bs, s, b.g, f, b.type, mass, aux, (C*)b.visc, b.newIf, b.donor, sc, k,
                                                      b.changed, b.changedCap);
    launch_gs<&k_fs_from_emptied<S>, 256, 1>(grid, bs, s, b.g, b.type, mass, aux, b.emptied, b.donor, b.newIf, b.ifCap, sc);
    launch_gs<&k_fs_update_mutants<S>, 256, 1>(grid, bs, s, b.type, mass, aux, b.filled, b.emptied, sc);
    launch_gs<&k_fs_isolated_a<S>, 256, 1>(grid, bs, s, b.g, b.type, mass, aux, b.ifList, b.newIf, b.donor, b.ifList2,
                                            b.ifCap, count2, sc);
    launch_gs<&k_fs_isolated_b<S>, 256, 1>(grid, bs, s, b.g, b.type, mass, b.ifList2, count2, sc, b.changed, b.changedCap);
    launch_gs<&k_fs_redistribute<S>, 256, 1>(grid, bs, s, b.type, mass, b.ifList2, count2, b.donor, b.filled, b.newIf, sc);
}

void lb_free_surface(Box& b, const Consts& k, DevScalars* sc, hipStream_t s) {
    if (b.prec == LBDEM_PREC_FP64) fs_impl<double>(b, k, sc, s);
    else if (b.prec == LBDEM_PREC_FP32) fs_impl<float>(b, k, sc, s);
    else fs_impl<__half>(b, k, sc, s);
}

// ------------------------------------------------------------------------------------------
// particle flags (LB.cpp:683-711 via 1875-1899; findNewActive LB.cpp:1980-2029)
// ------------------------------------------------------------------------------------------
// One wavefront per particle, over the particle's bounding box: active nodes inside the sphere,
// highest index wins. A node inside particle p and inside a higher-index particle q has
// |c_p - c_q| < r_p + r_q, so q is on p's Verlet list, which this rebuild step built from positions
// one step earlier with a skin of two length units (DEM.cpp:585-627). So p first gathers the
// higher-index partners whose spheres overlap its own (one list entry per lane), then writes every
// node inside its sphere and inside none of them with a plain store: each node gets exactly one
// writer, the highest index (the atomic maximum the list-free path uses gives the same). Without a
// list (no link cell along z, where the reference builds none) the atomic maximum is used.
// Particles are taken in link-cell order when the rebuild sorted them (neighbouring particles
// then share the type lines they read), in one contiguous share of that order per XCD (blocks
// b, b + 8, ... run on one XCD, whose L2 then holds one region). A wavefront walks its box 64 nodes
// per lane step, FLAGS_UNROLL steps at a time so that their type loads are in flight together; the
// sphere test keeps the reference's operations (LB.cpp:705, myvector.cpp:171-175).
constexpr int FLAGS_GRID = 2048;   // blocks of the flag pass (grid-stride; most steps return early)
constexpr int FLAGS_WAVES = 4;     // wavefronts (particles at a time) per block
constexpr int FLAGS_UNROLL = 6;
__device__ inline void k_flags_compute(Geom g, const uint8_t* type, const PartLB* plb, int nPart, int32_t* flagNew,
                                DevScalars* sc, const int* nbrCount, const int* nbrList, int useList,
                                const int* order) {
    // a coupling step that follows a coupling step without an LB step between them (the DEM
    // start-up repeat) finishes the previous findNewActive here (no other block touches kminP)
    if (blockIdx.x == 0 && threadIdx.x == 0) fna_finish(sc);
    if (!sc->newNeighborList) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) sc->kminP = INT_MAX;  // lowest flagged index, from k_flags_apply
    __shared__ double sq[FLAGS_WAVES][NBR_CAP][4];  // centre and radius of overlapping higher-index partners
    const int wv = __builtin_amdgcn_readfirstlane(int(threadIdx.x) >> 6), lane = int(threadIdx.x) & 63;
    double(*hq)[4] = sq[wv];
    const int dims[3] = {g.X, g.Y, g.Z};
    // this wavefront's particles: positions k0, k0 + step, ... below k1 of the (sorted) order
    int k0 = int(blockIdx.x) * FLAGS_WAVES + wv, k1 = nPart, step = int(gridDim.x) * FLAGS_WAVES;
    if (gridDim.x % 8 == 0) {
        const int share = (nPart + 7) / 8, xcd = int(blockIdx.x) % 8;
        k0 = xcd * share + (int(blockIdx.x) / 8) * FLAGS_WAVES + wv;
        k1 = min(nPart, (xcd + 1) * share);
        step = int(gridDim.x / 8) * FLAGS_WAVES;
    }
    for (int k = k0; k < k1; k += step) {
        const int p = order ? order[k] : k;
        const PartLB P = plb[p];
        int nHigher = 0;
        if (useList) {
            const int cnt = nbrCount[p];  // at most NBR_CAP - 1 entries: one lane each
            bool keep = false;
            double qc[4] = {0.0, 0.0, 0.0, 0.0};
            if (lane < cnt) {
                const int q = nbrList[p * NBR_CAP + lane];
                if (q > p) {
                    const PartLB Qp = plb[q];
                    const double dx = Qp.c[0] - P.c[0], dy = Qp.c[1] - P.c[1], dz = Qp.c[2] - P.c[2];
                    const double rs = P.r + Qp.r + 1e-9;  // a margin: a spare candidate only costs a test
                    keep = dx * dx + dy * dy + dz * dz < rs * rs;
                    qc[0] = Qp.c[0]; qc[1] = Qp.c[1]; qc[2] = Qp.c[2]; qc[3] = Qp.r;
                }
            }
            const uint64_t m = __ballot(keep);
            // the previous particle's reads of hq are done (LDS requests of a wavefront complete in order)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            if (keep) {
                const int slot = __popcll(m & ((uint64_t(1) << lane) - 1));
                hq[slot][0] = qc[0]; hq[slot][1] = qc[1]; hq[slot][2] = qc[2]; hq[slot][3] = qc[3];
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            nHigher = __popcll(m);
        }
        int lo[3], hi[3];
        for (int ax = 0; ax < 3; ++ax) {
            lo[ax] = int(floor(P.c[ax] - P.r)) - 1;
            hi[ax] = int(ceil(P.c[ax] + P.r)) + 1;
            lo[ax] = lo[ax] < 0 ? 0 : lo[ax];
            hi[ax] = hi[ax] > dims[ax] - 1 ? dims[ax] - 1 : hi[ax];
        }
        if (lo[0] > hi[0] || lo[1] > hi[1] || lo[2] > hi[2]) continue;
        const int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1, nxy = nx * ny, tot = nxy * (hi[2] - lo[2] + 1);
        // node c = lane + 64 j of the box, as (x, y, z) offsets advanced by 64 nodes per step
        const int sx = 64 % nx, sy = (64 / nx) % ny, sz = 64 / nxy;
        int ox = lane % nx, oy = (lane / nx) % ny, oz = lane / nxy;
        for (int c0 = 0; c0 < tot; c0 += 64 * FLAGS_UNROLL) {
            int64_t idx[FLAGS_UNROLL];
            bool in[FLAGS_UNROLL];
            uint8_t t[FLAGS_UNROLL];
#pragma unroll
            for (int u = 0; u < FLAGS_UNROLL; ++u) {
                const int x = lo[0] + ox, y = lo[1] + oy, z = lo[2] + oz;
                const bool ok = c0 + 64 * u + lane < tot;
                idx[u] = cidx(g, ok ? x : lo[0], ok ? y : lo[1], ok ? z : lo[2]);
                t[u] = type[idx[u]];
                in[u] = ok && inside_sphere(x, y, z, P.c[0], P.c[1], P.c[2], P.r);
                if (in[u] && useList)
                    for (int j = 0; j < nHigher && in[u]; ++j)
                        in[u] = !inside_sphere(x, y, z, hq[j][0], hq[j][1], hq[j][2], hq[j][3]);
                ox += sx;
                const int cy = ox >= nx ? 1 : 0;
                ox -= cy * nx;
                oy += sy + cy;
                const int cz = oy >= ny ? 1 : 0;
                oy -= cz * ny;
                oz += sz + cz;
            }
#pragma unroll
            for (int u = 0; u < FLAGS_UNROLL; ++u) {
                if (!in[u] || !is_active(t[u])) continue;
                if (useList) flagNew[idx[u]] = p;
                else atomicMax(&flagNew[idx[u]], p);
            }
        }
    }
}

void lb_flags_rebuild_compute(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, const int* nbrCount,
                              const int* nbrList, bool useList, const int* order, hipStream_t s) {
    // grid-stride over the particles: most steps only test the flag and return
    if (k.nPart > 0)
        launch_gs<&k_flags_compute, 64 * FLAGS_WAVES, 8, 16384>(
            std::min((k.nPart + FLAGS_WAVES - 1) / FLAGS_WAVES, FLAGS_GRID), 64 * FLAGS_WAVES, s, b.g, b.type, plb,
            k.nPart, b.flagNew, sc, nbrCount, nbrList, useList ? 1 : 0, order);
}

// clear every flag and solid index, apply the new flags (initializeParticleBoundaries)
__device__ inline void k_flags_apply(int64_t N, uint8_t* type, int32_t* solid, int32_t* flagNew, DevScalars* sc,
                                     int32_t* changed, int changedCap) {
    if (!sc->newNeighborList) return;
    __shared__ int bmin;
    if (threadIdx.x == 0) bmin = INT_MAX;
    __syncthreads();
    int mymin = INT_MAX;
    bool remark = false;  // a cleared wall index: link classes need a full pass
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x) {
        const int32_t s = flagNew[i];
        const uint8_t t = type[i];
        // an object node's index is part of its neighbours' link classes: re-mark around it
        // (a node covered by an impactor after the flags were computed takes a particle index)
        const bool obj = (t & T_BASE) == T_OBJ;
        if (s >= 0) {
            if (obj && solid[i] != s) push_changed(changed, changedCap, sc, int32_t(i));
            type[i] = uint8_t(t | T_PMASK);  // flagged nodes were active: the edge bit stays
            solid[i] = s;
            flagNew[i] = -1;
            mymin = min(mymin, int(i));
        } else {
            // active nodes keep their edge bit; other nodes lose the flag and a wall index
            const uint8_t nt = is_active(t) ? uint8_t(t & ~T_PMASK) : uint8_t(t & T_BASE);
            if (!is_active(t) && (t & 0x70)) remark = true;
            if (obj && solid[i] != 0) push_changed(changed, changedCap, sc, int32_t(i));
            if (nt != t) type[i] = nt;
            solid[i] = 0;
        }
    }
    if (__syncthreads_or(remark) && threadIdx.x == 0) atomicOr(&sc->fullMark, 1);
    if (mymin != INT_MAX) atomicMin(&bmin, mymin);
    __syncthreads();
    // blocks run roughly in index order: only a smaller minimum needs the global atomic
    if (threadIdx.x == 0 && bmin != INT_MAX &&
        bmin < __hip_atomic_load(&sc->kminP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(&sc->kminP, bmin);
}


void lb_flags_rebuild_apply(const Box& b, DevScalars* sc, hipStream_t s) {
    const int64_t nb = (b.g.N + 255) / 256;
    const int grid = int(nb < FLAGS_GRID ? nb : FLAGS_GRID);
    launch_gs<&k_flags_apply, 1024, 1, 16384>(grid, 256, s, b.g.N, b.type, b.solid, b.flagNew, sc, b.changed,
                                              b.changedCap);
}

__device__ inline void k_fna_check(Geom g, const int32_t* solid, const PartLB* plb, int nPart, DevScalars* sc) {
    fna_finish(sc);  // a scan still pending (boxes without particles have no k_flags_compute)
    if (sc->newNeighborList) {  // end of the flag pass (k_flags_compute set kminP to INT_MAX)
        if (sc->kminP == INT_MAX) sc->kminP = -1;
        sc->newNeighborList = 0;
    }
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

__device__ inline void k_fna_scan(Geom g, const uint8_t* type, const int32_t* solid, const PartLB* plb, int nPart,
                           DevScalars* sc) {
    if (!sc->needScan) return;
    __shared__ int bmin;
    if (threadIdx.x == 0) bmin = INT_MAX;
    __syncthreads();
    int mymin = INT_MAX;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        if (!(type[i] & T_PMASK)) continue;
        const int sp = solid[i];
        if (sp < 0 || sp >= nPart) continue;
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        const PartLB P = plb[sp];
        if (inside_sphere(x, y, z, P.c[0], P.c[1], P.c[2], P.r)) mymin = min(mymin, int(i));
    }
    if (mymin != INT_MAX) atomicMin(&bmin, mymin);
    __syncthreads();
    if (threadIdx.x == 0 && bmin != INT_MAX &&
        bmin < __hip_atomic_load(&sc->kstar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
        atomicMin(&sc->kstar, bmin);
}

__device__ inline void k_fna_unflag(Geom g, uint8_t* type, int32_t* solid, DevScalars* sc) {
    if (!sc->needScan) return;
    const int ks = sc->kstar;
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N && i < ks; i += int64_t(gridDim.x) * blockDim.x) {
        if (type[i] & T_PMASK) { type[i] &= uint8_t(~T_PMASK); solid[i] = 0; }
    }
}

void lb_find_new_active(const Box& b, const Consts& k, const PartLB* plb, DevScalars* sc, hipStream_t s) {
    launch<&k_fna_check, 1024, 1>(1, 1, s, b.g, b.solid, plb, k.nPart, sc);
    const int64_t nb = (b.g.N + 255) / 256;
    const int grid = int(nb < FLAGS_GRID ? nb : FLAGS_GRID);  // grid-stride; most steps return early
    launch_gs<&k_fna_scan, 1024, 1>(grid, 256, s, b.g, b.type, b.solid, plb, k.nPart, sc);
    launch_gs<&k_fna_unflag, 1024, 1>(grid, 256, s, b.g, b.type, b.solid, sc);
    // the scan's result (fna_finish) is taken by the next single-thread pass: k_begin_lb, or the
    // next coupling step's k_flags_compute
}

// ------------------------------------------------------------------------------------------
// impactor nodes (updateObjectBoundaries + findNewActiveObject, LB.cpp:797-943)
// ------------------------------------------------------------------------------------------
// uncovered object nodes (LB.cpp:846-943). The previous object nodes (ascending list) that the
// moved sphere no longer covers become fluid, interface or gas nodes. Pass 1 (k_obj_uncover_find,
// many blocks) compacts them in ascending order, each block over its slice of the list. Each
// conversion reads the node types its predecessors in that order wrote, so pass 2
// (k_obj_uncover, one block) takes the candidates in chunks of UNC_CHUNK: it stages their neighbour types in LDS, finds
// neighbours that are candidates of the chunk through a hash table, decides the new types (see
// below) and writes a chunk's results back before the next chunk stages its neighbourhood.
constexpr int UNCOVER_THREADS = 1024, UNC_CHUNK = 1024;
constexpr int UNC_MAX_BLOCKS = UNCOVER_BLOCKS_MAX, UNC_NODES_PER_BLOCK = 4096;
constexpr int UNC_HASH = 2048;  // open-addressing table of a chunk's candidates (at most half full)
__device__ inline uint32_t unc_hash(int32_t key) { return (uint32_t(key) * 2654435761u) >> (32 - 11); }
__device__ inline uint8_t uncover_type(bool anyFluid, bool anyGas) {
    return !anyFluid ? uint8_t(T_GAS) : (anyGas ? uint8_t(T_IF) : uint8_t(T_FLUID));
}
__device__ inline int block_exclusive_scan(int v, int* sWave, int& total) {
    const int lane = __lane_id(), wave = threadIdx.x / 64, nw = blockDim.x / 64;
    int incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d, 64);
        if (lane >= d) incl += o;
    }
    if (lane == 63) sWave[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int w = 0; w < nw; ++w) { const int c = sWave[w]; sWave[w] = acc; acc += c; }
        sWave[nw] = acc;
    }
    __syncthreads();
    const int r = sWave[wave] + incl - v;
    total = sWave[nw];
    __syncthreads();
    return r;
}

// pass 1: ordered compaction of the nodes of each block's slice of the list that left the sphere,
// into the same slice of cand
__device__ inline void k_obj_uncover_find(Geom g, const int32_t* objList, int32_t* cand, int* blockCount, int G,
                                          const DevScalars* sc, const ObjLB* olb) {
    if (int(blockIdx.x) >= G) return;  // G: the box's grid
    __shared__ int sWave[UNCOVER_THREADS / 64 + 1];
    const int tid = threadIdx.x;
    const int n = sc->nObjList;
    const int per = (n + G - 1) / G;
    const ObjLB O = olb[0];
    const int beg = min(n, int(blockIdx.x) * per), end = min(n, beg + per);
    int m = 0;
    for (int base = beg; base < end; base += UNCOVER_THREADS) {
        const int kk = base + tid;
        int32_t i = 0;
        int flag = 0;
        if (kk < end) {
            i = objList[kk];
            int x, y, z;
            cell_xyz(g, i, x, y, z);
            flag = !inside_sphere(x, y, z, O.c[0], O.c[1], O.c[2], O.r);
        }
        int tot;
        const int pos = block_exclusive_scan(flag, sWave, tot);
        if (flag) cand[beg + m + pos] = i;
        m += tot;
    }
    if (tid == 0) blockCount[blockIdx.x] = m;
}

// pass 2 (one block): the candidates of the G slices in order
template <typename S, int L>
__device__ inline void k_obj_uncover(Geom g, S* f, uint8_t* type,
                                                                 typename Types<S>::M* mass,
                                                                 typename Types<S>::C* aux, typename Types<S>::C* visc,
                                                                 const int32_t* cand, const int* blockCount, int G,
                                                                 DevScalars* sc, Consts k, int32_t* changed, int changedCap) {
    using C = typename Types<S>::C;
    __shared__ int sWave[UNCOVER_THREADS / 64 + 1];
    __shared__ uint8_t nbT[Q - 1][UNC_CHUNK];   // neighbour base types (values before this pass)
    __shared__ uint8_t newT[UNC_CHUNK];
    __shared__ int32_t sCand[UNC_CHUNK];
    __shared__ int32_t hKey[UNC_HASH];          // chunk candidates: node index -> position
    __shared__ int16_t hPos[UNC_HASH];
    __shared__ int sPre[UNC_MAX_BLOCKS + 1];   // exclusive prefix of the slices' candidate counts
    __shared__ uint8_t ownT[UNC_CHUNK];        // the candidates' own types (staged with the neighbours')
    __shared__ int sChBase;                    // the chunk's first slot on the change list
    const int tid = threadIdx.x;
    const int n = sc->nObjList;
    const int per = (n + G - 1) / G;
    {
        const int c = tid < G ? blockCount[tid] : 0;
        int tot;
        const int ex = block_exclusive_scan(c, sWave, tot);
        if (tid < G) sPre[tid] = ex;
        if (tid == 0) sPre[G] = tot;
    }
    __syncthreads();
    const int m = sPre[G];
    // pass 2: candidates in ascending node order, in chunks; a candidate's new type depends on
    // the new types of earlier candidates among its neighbours (the reference's sequential order)
    for (int c0 = 0; c0 < m; c0 += UNC_CHUNK) {
        const int cnt = min(UNC_CHUNK, m - c0);
        if (tid < cnt) {
            const int q = c0 + tid;
            int lo = 0, hi = G - 1;  // block b with sPre[b] <= q < sPre[b + 1]
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (sPre[mid] <= q) lo = mid; else hi = mid - 1;
            }
            sCand[tid] = cand[min(n, lo * per) + (q - sPre[lo])];
        }
        for (int h = tid; h < UNC_HASH; h += UNCOVER_THREADS) hKey[h] = -1;
        __syncthreads();
        if (tid < cnt) {  // position of every candidate, by node index
            const int32_t key = sCand[tid];
            for (uint32_t h = unc_hash(key);; h = (h + 1) & (UNC_HASH - 1))
                if (atomicCAS(&hKey[h], -1, key) == -1) { hPos[h] = int16_t(tid); break; }
        }
        if (tid == 0) sChBase = atomicAdd(&sc->nChanged, cnt);  // every candidate goes on the change list
        if (tid < cnt) {
            const int32_t i = sCand[tid];
            int x, y, z;
            cell_xyz(g, i, x, y, z);
            uint8_t nt[Q];  // the 19 loads are issued together, then written to LDS
            nt[0] = type[i];
#pragma unroll
            for (int j = 1; j < Q; ++j) nt[j] = type[nbr_plain(g, x, y, z, j)] & T_BASE;
            ownT[tid] = nt[0];
#pragma unroll
            for (int j = 1; j < Q; ++j) nbT[j - 1][tid] = nt[j];
        }
        __syncthreads();
        // The new type of candidate t: gas without a fluid neighbour, interface with fluid and gas
        // neighbours, fluid otherwise, over the types its neighbours have when the reference
        // reaches t (earlier candidates converted, later ones still object nodes). One wavefront
        // takes 64 candidates at a time: types of earlier groups are final; dependencies inside
        // the group are resolved by repeating the decision with the group's current types until
        // none changes. On the dependency order (a candidate depends on earlier ones only) the
        // only fixed point is the sequential result, reached after at most depth + 1 rounds.
        if (tid < 64) {
            const int lane = tid;
            for (int s0 = 0; s0 < cnt; s0 += 64) {
                const int t = s0 + lane;
                const bool valid = t < cnt;
                bool baseF = false, baseG = false;
                uint64_t dep = 0;
                if (valid) {
                    int x, y, z;
                    cell_xyz(g, sCand[t], x, y, z);
                    for (int j = 1; j < Q; ++j) {
                        const int32_t nb = int32_t(nbr_plain(g, x, y, z, j));
                        int pos = -1;
                        for (uint32_t h = unc_hash(nb);; h = (h + 1) & (UNC_HASH - 1)) {
                            const int32_t kk = hKey[h];
                            if (kk == nb) { pos = hPos[h]; break; }
                            if (kk == -1) break;
                        }
                        uint8_t ty = nbT[j - 1][t];
                        if (pos >= 0 && pos < t) {
                            if (pos < s0) ty = newT[pos];
                            else { dep |= 1ull << (pos - s0); continue; }
                        }
                        baseF |= ty == T_FLUID;
                        baseG |= ty == T_GAS;
                    }
                }
                uint8_t myT = uncover_type(baseF, baseG);
                for (int round = 0; round <= 64; ++round) {
                    const uint64_t fm = __ballot(valid && myT == T_FLUID);
                    const uint64_t gm = __ballot(valid && myT == T_GAS);
                    const uint8_t nt = uncover_type(baseF || (dep & fm), baseG || (dep & gm));
                    const uint64_t ch = __ballot(valid && nt != myT);
                    myT = nt;
                    if (!ch) break;
                }
                if (valid) newT[t] = myT;
                __threadfence_block();  // the next group of this wavefront reads these
            }
        }
        __syncthreads();
        if (tid < cnt) {
            const int32_t i = sCand[tid];
            const uint8_t nt = newT[tid];
            type[i] = uint8_t((ownT[tid] & T_PMASK) | nt);  // no other thread writes a candidate's type
            const int pos = sChBase + tid;
            if (pos < changedCap) changed[pos] = i;
            else atomicOr(&sc->fullMark, 1);
            if (nt != T_GAS) {
                const bool isIf = nt == T_IF;
                int x, y, z;
                cell_xyz(g, i, x, y, z);
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
                if (visc) visc[i] = C(k.visc);  // initDynVisc (LB.cpp:928-934)
            }
        }
        __threadfence_block();
        __syncthreads();
    }
}

// cover: every node inside an object sphere becomes an object node; ascending object node list.
// Pass a covers the union box of the spheres and counts per block (blocks own consecutive
// ranges of the box); pass b writes the list at the blocks' exclusive prefix offsets.
constexpr int COVER_GRID = COVER_BLOCKS, COVER_BS = 256;
struct CoverBox {
    int lo[3], n[3], tot;
};
__device__ inline CoverBox cover_box(const Geom& g, const ObjLB* olb, int nObj) {
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    const int dims[3] = {g.X, g.Y, g.Z};
    for (int o = 0; o < nObj; ++o)
        for (int a = 0; a < 3; ++a) {
            const int l = int(floor(olb[o].c[a] - olb[o].r)) - 1, h = int(ceil(olb[o].c[a] + olb[o].r)) + 1;
            lo[a] = min(lo[a], max(l, 0));
            hi[a] = max(hi[a], min(h, dims[a] - 1));
        }
    CoverBox b;
    for (int a = 0; a < 3; ++a) { b.lo[a] = lo[a]; b.n[a] = max(hi[a] - lo[a] + 1, 0); }
    b.tot = b.n[0] * b.n[1] * b.n[2];
    return b;
}
__device__ inline int inside_which(int x, int y, int z, const ObjLB* olb, int nObj) {
    int which = -1;
    for (int o = 0; o < nObj; ++o)
        if (inside_sphere(x, y, z, olb[o].c[0], olb[o].c[1], olb[o].c[2], olb[o].r)) which = o;
    return which;
}
__device__ inline void k_obj_cover_a(Geom g, uint8_t* type, int32_t* solid, const ObjLB* olb,
                                                          int nObj, int* blockCount, int32_t* seg, int segCap,
                                                          int32_t* changed, int changedCap, DevScalars* sc) {
    __shared__ int sWave[COVER_BS / 64 + 1];
    const CoverBox cb = cover_box(g, olb, nObj);
    const int per = (cb.tot + gridDim.x - 1) / gridDim.x;
    const int beg = min(cb.tot, int(blockIdx.x) * per), end = min(cb.tot, beg + per);
    const bool segmented = cb.tot <= segCap;  // the block's nodes go to seg[beg ...] in order
    int cnt = 0;
    for (int c0 = beg; c0 < end; c0 += COVER_BS) {
        const int c = c0 + threadIdx.x;
        int in = 0;
        int64_t i = 0;
        if (c < end) {
            const int x = cb.lo[0] + c % cb.n[0], y = cb.lo[1] + (c / cb.n[0]) % cb.n[1], z = cb.lo[2] + c / (cb.n[0] * cb.n[1]);
            const int which = inside_which(x, y, z, olb, nObj);
            if (which >= 0) {
                i = cidx(g, x, y, z);
                const uint8_t t = type[i];
                // the link classes of neighbouring edge nodes hold the type and the object index
                if ((t & T_BASE) != T_OBJ || solid[i] != which) push_changed(changed, changedCap, sc, int32_t(i));
                type[i] = uint8_t((t & T_PMASK) | T_OBJ);
                solid[i] = which;
                in = 1;
            }
        }
        if (segmented) {
            int tot;
            const int pos = block_exclusive_scan(in, sWave, tot);
            if (in) seg[beg + cnt + pos] = int32_t(i);
            cnt += tot;
        } else {
            cnt += __syncthreads_count(in);
        }
    }
    if (threadIdx.x == 0) blockCount[blockIdx.x] = cnt;
}
// the object node list: every block's nodes at the exclusive prefix of the block counts
__device__ inline void k_obj_cover_b(Geom g, const uint8_t* type, const ObjLB* olb, int nObj,
                                                          const int* blockCount, const int32_t* seg, int segCap,
                                                          int32_t* objList, int cap, DevScalars* sc) {
    __shared__ int sWave[COVER_BS / 64 + 1];
    __shared__ int sOff, sTot;
    const CoverBox cb = cover_box(g, olb, nObj);
    {
        int before = 0, all = 0;
        for (int b = threadIdx.x; b < int(gridDim.x); b += COVER_BS) {
            const int v = blockCount[b];
            all += v;
            if (b < int(blockIdx.x)) before += v;
        }
        int t1, t2;
        (void)block_exclusive_scan(before, sWave, t1);
        (void)block_exclusive_scan(all, sWave, t2);
        if (threadIdx.x == 0) { sOff = t1; sTot = t2; }
        __syncthreads();
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->nObjList = sTot < cap ? sTot : cap;
        if (sTot > cap) sc->listOverflow |= 1;
    }
    int off = sOff;
    const int per = (cb.tot + gridDim.x - 1) / gridDim.x;
    const int beg = min(cb.tot, int(blockIdx.x) * per), end = min(cb.tot, beg + per);
    if (cb.tot <= segCap) {
        const int cnt = blockCount[blockIdx.x];
        for (int k = threadIdx.x; k < cnt; k += COVER_BS)
            if (off + k < cap) objList[off + k] = seg[beg + k];
        return;
    }
    // union box larger than the staging buffer (impactors far apart): scan the box again
    for (int c0 = beg; c0 < end; c0 += COVER_BS) {
        const int c = c0 + threadIdx.x;
        int in = 0;
        int64_t i = 0;
        if (c < end) {
            const int x = cb.lo[0] + c % cb.n[0], y = cb.lo[1] + (c / cb.n[0]) % cb.n[1], z = cb.lo[2] + c / (cb.n[0] * cb.n[1]);
            i = cidx(g, x, y, z);
            in = (type[i] & T_BASE) == T_OBJ && inside_which(x, y, z, olb, nObj) >= 0;
        }
        int tot;
        const int pos = block_exclusive_scan(in, sWave, tot);
        if (in && off + pos < cap) objList[off + pos] = int32_t(i);
        off += tot;
    }
}

template <typename S>
static void obj_impl(Box& b, const Consts& k, const ObjLB* olb, DevScalars* sc, hipStream_t s) {
    using M = typename Types<S>::M;
    using C = typename Types<S>::C;
    // the object node list holds at most objCap nodes: one block per UNC_NODES_PER_BLOCK of them
    const int grid = std::max(1, std::min(UNC_MAX_BLOCKS, (b.objCap + UNC_NODES_PER_BLOCK - 1) / UNC_NODES_PER_BLOCK));
    int* uncCount = b.coverCount + COVER_GRID;
    launch<&k_obj_uncover_find, UNCOVER_THREADS, 1>(grid, UNCOVER_THREADS, s, b.g, b.objList, b.objCand, uncCount, grid, sc, olb);
    if (b.parity == 0)
        launch<&k_obj_uncover<S, 0>, UNCOVER_THREADS, 1>(1, UNCOVER_THREADS, s, b.g, (S*)b.f, b.type, (M*)b.mass, (C*)b.aux,
                                                         (C*)b.visc, b.objCand, uncCount, grid, sc, k, b.changed, b.changedCap);
    else
        launch<&k_obj_uncover<S, 1>, UNCOVER_THREADS, 1>(1, UNCOVER_THREADS, s, b.g, (S*)b.f, b.type, (M*)b.mass, (C*)b.aux,
                                                         (C*)b.visc, b.objCand, uncCount, grid, sc, k, b.changed, b.changedCap);
    // the candidate buffer of the uncover pass stages the cover pass's per-block node lists
    launch_gs<&k_obj_cover_a, COVER_BS, 1>(COVER_GRID, COVER_BS, s, b.g, b.type, b.solid, olb, k.nObj, b.coverCount, b.objCand, b.objCap, b.changed, b.changedCap, sc);
    launch_gs<&k_obj_cover_b, COVER_BS, 1>(COVER_GRID, COVER_BS, s, b.g, b.type, olb, k.nObj, b.coverCount, b.objCand, b.objCap, b.objList, b.objCap, sc);
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
// post-streaming populations q0 <= q < q1 of every node in FP64 (plane q at out + (q - q0) N),
// zero where the node is not active
template <typename S>
__device__ inline void k_download_f(Geom g, const S* f, const uint8_t* type, int L, int q0, int q1, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        const bool act = is_active(type[i]);
        int x, y, z;
        cell_xyz(g, i, x, y, z);
        if (!act) { for (int q = q0; q < q1; ++q) out[int64_t(q - q0) * g.N + i] = 0.0; continue; }
        const Offs o(g, x, y, z);
        for (int q = q0; q < q1; ++q)
            out[int64_t(q - q0) * g.N + i] = wq(q) + double(StoreT<S>::ld(f + load_slot(g, i, o, q, L)));
    }
}
void lb_download_f(const Box& b, double* out, hipStream_t s, int q0, int q1) {
    const int grid = 4096;
    if (b.prec == LBDEM_PREC_FP64) launch<&k_download_f<double>, 1024, 1>(grid, 256, s, b.g, (const double*)b.f, b.type, b.parity, q0, q1, out);
    else if (b.prec == LBDEM_PREC_FP32) launch<&k_download_f<float>, 1024, 1>(grid, 256, s, b.g, (const float*)b.f, b.type, b.parity, q0, q1, out);
    else launch<&k_download_f<__half>, 1024, 1>(grid, 256, s, b.g, (const __half*)b.f, b.type, b.parity, q0, q1, out);
}

// mass field in FP64, zero where the node is not active
template <typename M>
__device__ inline void k_download_mass(int64_t N, const uint8_t* type, const M* mass, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < N; i += int64_t(gridDim.x) * blockDim.x)
        out[i] = is_active(type[i]) ? double(mass[i]) : 0.0;
}
void lb_download_mass(const Box& b, double* out, hipStream_t s) {
    if (b.prec == LBDEM_PREC_FP64) launch<&k_download_mass<double>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const double*)b.mass, out);
    else launch<&k_download_mass<float>, 1024, 1>(4096, 256, s, b.g.N, b.type, (const float*)b.mass, out);
}

template <typename S>
__device__ inline void k_download_n(Geom g, const uint8_t* type, const typename Types<S>::M* mass,
                             const typename Types<S>::C* aux, double* out) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < g.N; i += int64_t(gridDim.x) * blockDim.x) {
        const uint8_t t = type[i] & T_BASE;
        out[i] = t == T_FLUID ? double(mass[i]) : (t == T_IF ? double(aux[i]) : 0.0);
    }
}
void lb_download_n(const Box& b, const Consts& k, double* out, hipStream_t s) {
    (void)k;
    if (b.prec == LBDEM_PREC_FP64) launch<&k_download_n<double>, 1024, 1>(4096, 256, s, b.g, b.type, (const double*)b.mass, (const double*)b.aux, out);
    else launch<&k_download_n<float>, 1024, 1>(4096, 256, s, b.g, b.type, (const float*)b.mass, (const float*)b.aux, out);
}

// total fluid mass: active nodes outside particles (LB.cpp:388-395)
template <typename M>
__device__ inline void k_total_mass(int64_t N, const uint8_t* type, const M* mass, double* out) {
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
    if (b.prec == LBDEM_PREC_FP64) launch<&k_total_mass<double>, 1024, 1>(1024, 256, s, b.g.N, b.type, (const double*)b.mass, out);
    else launch<&k_total_mass<float>, 1024, 1>(1024, 256, s, b.g.N, b.type, (const float*)b.mass, out);
}

__device__ inline void k_count_types(int64_t N, const uint8_t* type, int* counts) {
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
    launch<&k_count_types, 1024, 1>(1024, 256, s, b.g.N, b.type, counts);
}

}  // namespace lbdem
