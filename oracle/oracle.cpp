// ORACLE — TEST INFRASTRUCTURE ONLY. Nothing in the product links, loads or calls
// this file. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// use it, and only as the checker (or as the timed CPU baseline).
//
// CPU restatement of the VISCOPLASTIC-LBM time step, following the reference's
// own data layout and order of operations so that the GPU product (which uses
// a different, pull-streaming design) can be checked against it:
//
//   * populations: direction-major SoA with one ghost z-plane,
//       idx = x + NX*(y + NY*(z + (NZ+1)*i))        (src/CUDA/globalFunctions.h:143-147)
//   * one step = collide+push-stream every used node into popAux, storing the
//     post-collision populations back into pop at BC nodes
//                                                   (src/CUDA/lbm.cu:3-504)
//     -> fold the ghost plane z=NZ into z=0 / z=NZ-1 (src/CUDA/lbm.cu:600-657)
//     -> apply the BC list in z,y,x order            (src/CUDA/lbm.cu:579-598,
//                                                    src/CUDA/structs/boundaryConditionsInfo.h:76-116)
//     -> swap pop/popAux                             (src/CUDA/structs/populations.h:64-70)
//   * node-type map words, 32 bits per node        (src/CUDA/structs/nodeTypeMap.h:15-322)
//   * box builders as priority if-chains            (src/CUDA/boundaryConditionsBuilders/*.cu)
//
// Parity status of each function is stated at its definition ("pinned by ...").
// Arithmetic follows the reference expression by expression (same summation
// groupings, same recursive terms[] reuse, same double literals), templated on
// the storage type so the FP32 build keeps the reference's mixed-precision
// promotions.
//
// Built by oracle/Makefile into oracle/liboracle.so (g++ -O2 -fopenmp).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <chrono>
#include <omp.h>

namespace {

// ---------------------------------------------------------------------------
// Velocity sets (src/CUDA/velocitySets/D3Q19.h:31-46, D3Q27.h:39-56).
// Pinned by closed-form invariants in tests/test_oracle_pins.py.
// ---------------------------------------------------------------------------
const int CX[27] = {0, 1,-1, 0, 0, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1};
const int CY[27] = {0, 0, 0, 1,-1, 0, 0, 1,-1, 0, 0, 1,-1,-1, 1, 0, 0, 1,-1, 1,-1, 1,-1,-1, 1, 1,-1};
const int CZ[27] = {0, 0, 0, 0, 0, 1,-1, 0, 0, 1,-1, 1,-1, 0, 0,-1, 1,-1, 1, 1,-1,-1, 1, 1,-1, 1,-1};

// ---------------------------------------------------------------------------
// Node-type map bit layout (src/CUDA/structs/nodeTypeMap.h:15-108).
// ---------------------------------------------------------------------------
enum : uint32_t {
    OFF_INTERP = 23, OFF_SAVE = 22, OFF_USED = 21, OFF_SCHEME = 18, OFF_DIR = 13,
    OFF_GEOM = 12, OFF_UX = 9, OFF_UY = 6, OFF_UZ = 3, OFF_RHO = 0
};
enum Scheme : int { BC_NULL = 0, VEL_ZOUHE = 1, VEL_BB = 2, PRES_ZOUHE = 3, FREE_SLIP = 4,
                    BOUNCE_BACK = 5, INTERP_BB = 6, SPECIAL = 7 };
enum Dir : int { N = 0, S, W, E, F, B, NW, NE, NF, NB, SW, SE, SF, SB, WF, WB, EF, EB,
                 NWF, NWB, NEF, NEB, SWF, SWB, SEF, SEB };

// Setters restate the reference's guarded read-modify-write updates.
struct Word {
    uint32_t m = 0;
    void used(bool v) { if (v) m |= (1u << OFF_USED); else m &= ~(1u << OFF_USED); }
    void save(bool v) { if (v) m |= (1u << OFF_SAVE); else m &= ~(1u << OFF_SAVE); }
    void field(uint32_t off, uint32_t width_mask, uint32_t v) {
        if (v <= width_mask) m = (m & ~(width_mask << off)) | (v << off);
    }
    void scheme(int s) { field(OFF_SCHEME, 7u, (uint32_t)s); }
    void dir(int d) { field(OFF_DIR, 31u, (uint32_t)d); }
    void geom(int g) { field(OFF_GEOM, 1u, (uint32_t)g); }
    void uxi(int i) { field(OFF_UX, 7u, (uint32_t)i); }
    void uyi(int i) { field(OFF_UY, 7u, (uint32_t)i); }
    void uzi(int i) { field(OFF_UZ, 7u, (uint32_t)i); }
    void rhoi(int i) { field(OFF_RHO, 7u, (uint32_t)i); }
    void unknown_interp(int pop) { if (pop < 8) m |= ((1u << pop) << OFF_INTERP); }
};
inline bool w_used(uint32_t m) { return (m >> OFF_USED) & 1u; }
inline int w_scheme(uint32_t m) { return (m >> OFF_SCHEME) & 7u; }
inline int w_dir(uint32_t m) { return (m >> OFF_DIR) & 31u; }
inline int w_geom(uint32_t m) { return (m >> OFF_GEOM) & 1u; }
inline int w_uxi(uint32_t m) { return (m >> OFF_UX) & 7u; }
inline int w_uyi(uint32_t m) { return (m >> OFF_UY) & 7u; }
inline int w_uzi(uint32_t m) { return (m >> OFF_UZ) & 7u; }
inline int w_rhoi(uint32_t m) { return (m >> OFF_RHO) & 7u; }
inline unsigned w_interp_bits(uint32_t m) { return (m >> OFF_INTERP) & 0xffu; }
// getSavePostCol is true for an explicit save bit or any BC scheme
// (src/CUDA/structs/nodeTypeMap.h:158-163).
inline bool w_save_post_col(uint32_t m) { return ((m >> OFF_SAVE) & 1u) || (w_scheme(m) != BC_NULL); }

} // namespace

extern "C" {

// Descriptor shared with tests (ctypes mirrors this layout field by field).
struct or_desc {
    int nx, ny, nz, q;
    int rheo;              // 0 Newtonian, 1 Bingham, 2 power-law
    int builder;           // builder id, see or_build_map
    double tau, rho0, fx, fy, fz;
    double eta_p, s_y;     // Bingham (src/CUDA/NNF/nnf.h:18-24)
    double n_index, k_consistency, gamma_0;   // power-law (src/CUDA/NNF/nnf.h:10-16)
    int omega_clamp;       // extension: clamp tau into [tau_min, tau_max]
    double tau_min, tau_max;
    double ux_bc[8], uy_bc[8], uz_bc[8], rho_bc[8];   // src/CUDA/var.h:114-117
    double duct_outer_radius, duct_inner_radius;      // interpolated-BB builders
    double outer_rotation, outer_velocity, inner_rotation, inner_velocity;
    int collision;         // 0 regularized (lbm.cu:263-461), 1 BGK (recursiveCollisionForceNoReg.txt)
    double les_const;      // LES_MODEL Implicit_const (lbm.cu:226-242); 0 = no LES
    int les_visc_turb;     // LES_EXPORT_VISC_TURBULENT (lesVar.h:16): keep tau_t/3 per node
};

} // extern "C"

namespace {

// ---------------------------------------------------------------------------
// Builders. Each is a restatement of one reference builder's if-chain
// (corner -> edge -> face priority). Region order is the reference chain
// order; an empty action leaves the defaults (used, no save, BC_NULL,
// CONCAVE, value indices 0).
// Pinned by: tests check geometric descriptions from the reference file
// headers (e.g. "N, S: wall; B, F: periodic; W, E: periodic").
// ---------------------------------------------------------------------------
enum Region { R_SWB, R_SWF, R_SEB, R_SEF, R_NWB, R_NWF, R_NEB, R_NEF, R_SW, R_SE, R_NW, R_NE,
              R_SB, R_SF, R_NB, R_NF, R_WB, R_WF, R_EB, R_EF, R_S, R_N, R_W, R_E, R_B, R_F, R_INT };

int classify(int x, int y, int z, int nx, int ny, int nz) {
    const bool s = (y == 0), n = (y == ny - 1), w = (x == 0), e = (x == nx - 1);
    const bool b = (z == 0), f = (z == nz - 1);
    if (s && w && b) return R_SWB;
    else if (s && w && f) return R_SWF;
    else if (s && e && b) return R_SEB;
    else if (s && e && f) return R_SEF;
    else if (n && w && b) return R_NWB;
    else if (n && w && f) return R_NWF;
    else if (n && e && b) return R_NEB;
    else if (n && e && f) return R_NEF;
    else if (s && w) return R_SW;
    else if (s && e) return R_SE;
    else if (n && w) return R_NW;
    else if (n && e) return R_NE;
    else if (s && b) return R_SB;
    else if (s && f) return R_SF;
    else if (n && b) return R_NB;
    else if (n && f) return R_NF;
    else if (w && b) return R_WB;
    else if (w && f) return R_WF;
    else if (e && b) return R_EB;
    else if (e && f) return R_EF;
    else if (s) return R_S;
    else if (n) return R_N;
    else if (w) return R_W;
    else if (e) return R_E;
    else if (b) return R_B;
    else if (f) return R_F;
    return R_INT;
}

struct Act { int scheme, dir, uxi, uyi, uzi, rhoi; };
const Act NONE = {-1, 0, 0, 0, 0, 0};
inline Act bb(int d) { return {BOUNCE_BACK, d, 0, 0, 0, 0}; }

// One action per region, in Region order; last entry (interior) always NONE.
void builder_actions(int id, Act a[27]) {
    for (int r = 0; r < 27; r++) a[r] = NONE;
    switch (id) {
    case 0: // taylorGreenVortex.cu: fully periodic
        break;
    case 1: // parallelPlatesBounceBack_Y.cu:35-174 — walls N,S
        for (int r : {R_SWB, R_SWF, R_SEB, R_SEF, R_SW, R_SE, R_SB, R_SF, R_S}) a[r] = bb(S);
        for (int r : {R_NWB, R_NWF, R_NEB, R_NEF, R_NW, R_NE, R_NB, R_NF, R_N}) a[r] = bb(N);
        break;
    case 2: // parallelPlatesBounceBack_X.cu — walls W,E
        for (int r : {R_SWB, R_SWF, R_NWB, R_NWF, R_SW, R_NW, R_WB, R_WF, R_W}) a[r] = bb(W);
        for (int r : {R_SEB, R_SEF, R_NEB, R_NEF, R_SE, R_NE, R_EB, R_EF, R_E}) a[r] = bb(E);
        break;
    case 3: // parallelPlatesBounceBack_Z.cu — walls B,F
        for (int r : {R_SWB, R_SEB, R_NWB, R_NEB, R_SB, R_NB, R_WB, R_EB, R_B}) a[r] = bb(B);
        for (int r : {R_SWF, R_SEF, R_NWF, R_NEF, R_SF, R_NF, R_WF, R_EF, R_F}) a[r] = bb(F);
        break;
    case 4: // squareDuctBounceBack.cu:33-184 — walls N,S,W,E; periodic B,F
    case 11: // squareDuctZouHe (new): same walls, velocity Zou-He inlet B, pressure Zou-He outlet F
        a[R_SWB] = bb(SW); a[R_SWF] = bb(SW); a[R_SEB] = bb(SE); a[R_SEF] = bb(SE);
        a[R_NWB] = bb(NW); a[R_NWF] = bb(NW); a[R_NEB] = bb(NE); a[R_NEF] = bb(NE);
        a[R_SW] = bb(SW); a[R_SE] = bb(SE); a[R_NW] = bb(NW); a[R_NE] = bb(NE);
        a[R_SB] = bb(S); a[R_SF] = bb(S); a[R_NB] = bb(N); a[R_NF] = bb(N);
        a[R_WB] = bb(W); a[R_WF] = bb(W); a[R_EB] = bb(E); a[R_EF] = bb(E);
        a[R_S] = bb(S); a[R_N] = bb(N); a[R_W] = bb(W); a[R_E] = bb(E);
        if (id == 11) {
            a[R_B] = {VEL_ZOUHE, B, 0, 0, 1, 0};
            a[R_F] = {PRES_ZOUHE, F, 0, 0, 0, 0};
        }
        break;
    case 5: // closedBoxBounceBack.cu:33-186 — every region its own direction
        a[R_SWB] = bb(SWB); a[R_SWF] = bb(SWF); a[R_SEB] = bb(SEB); a[R_SEF] = bb(SEF);
        a[R_NWB] = bb(NWB); a[R_NWF] = bb(NWF); a[R_NEB] = bb(NEB); a[R_NEF] = bb(NEF);
        a[R_SW] = bb(SW); a[R_SE] = bb(SE); a[R_NW] = bb(NW); a[R_NE] = bb(NE);
        a[R_SB] = bb(SB); a[R_SF] = bb(SF); a[R_NB] = bb(NB); a[R_NF] = bb(NF);
        a[R_WB] = bb(WB); a[R_WF] = bb(WF); a[R_EB] = bb(EB); a[R_EF] = bb(EF);
        a[R_S] = bb(S); a[R_N] = bb(N); a[R_W] = bb(W); a[R_E] = bb(E);
        a[R_B] = bb(B); a[R_F] = bb(F);
        break;
    case 6: { // parallelPlatesZouHe.cu — velocity Zou-He N,S; pressure Zou-He B,F
        const Act vs = {VEL_ZOUHE, S, 0, 0, 0, 0}, vn = {VEL_ZOUHE, N, 0, 0, 0, 0};
        const Act pb = {PRES_ZOUHE, B, 0, 0, 0, 1}, pf = {PRES_ZOUHE, F, 0, 0, 0, 2};
        for (int r : {R_SWB, R_SWF, R_SEB, R_SEF, R_SW, R_SE, R_SB, R_SF, R_S}) a[r] = vs;
        for (int r : {R_NWB, R_NWF, R_NEB, R_NEF, R_NW, R_NE, R_NB, R_NF, R_N}) a[r] = vn;
        a[R_WB] = pb; a[R_WF] = pf; a[R_EB] = pb; a[R_EF] = pf;
        a[R_B] = pb; a[R_F] = pf;
        break; }
    case 7: { // lidDrivenBounceBack.cu — moving lid N (ux index 1)
        const Act lid = {VEL_BB, N, 1, 0, 0, 0};
        a[R_SWB] = bb(SW); a[R_SWF] = bb(SW); a[R_SEB] = bb(SE); a[R_SEF] = bb(SE);
        a[R_NWB] = bb(NW); a[R_NWF] = bb(NW); a[R_NEB] = bb(NE); a[R_NEF] = bb(NE);
        a[R_SW] = bb(SW); a[R_SE] = bb(SE); a[R_NW] = bb(NW); a[R_NE] = bb(NE);
        a[R_SB] = bb(S); a[R_SF] = bb(S); a[R_NB] = lid; a[R_NF] = lid;
        a[R_WB] = bb(W); a[R_WF] = bb(W); a[R_EB] = bb(E); a[R_EF] = bb(E);
        a[R_S] = bb(S); a[R_N] = lid; a[R_W] = bb(W); a[R_E] = bb(E);
        break; }
    default:
        break;
    }
}

// Builders whose chain tests fewer conditions than the full 26-region chain.
bool short_chain_word(int id, int x, int y, int z, int nx, int ny, int nz, Word& w) {
    (void)x; (void)nx;
    if (id == 8) { // couetteZouHe.cu: S (uz index 2), N (uz index 1)
        if (y == 0) { w.scheme(VEL_ZOUHE); w.dir(S); w.uzi(2); }
        else if (y == ny - 1) { w.scheme(VEL_ZOUHE); w.dir(N); w.uzi(1); }
        return true;
    }
    if (id == 9) { // inflowFreeSlip.cu: free slip N,S; velocity bounce-back B,F (uz index 1)
        const bool b = (z == 0), f = (z == nz - 1);
        if (y == 0 && b) { w.scheme(FREE_SLIP); w.dir(S); }
        else if (y == 0 && f) { w.scheme(FREE_SLIP); w.dir(S); }
        else if (y == ny - 1 && b) { w.scheme(FREE_SLIP); w.dir(N); }
        else if (y == ny - 1 && f) { w.scheme(FREE_SLIP); w.dir(N); }
        else if (y == 0) { w.scheme(FREE_SLIP); w.dir(S); }
        else if (y == ny - 1) { w.scheme(FREE_SLIP); w.dir(N); }
        else if (b) { w.scheme(VEL_BB); w.dir(B); w.uzi(1); }
        else if (f) { w.scheme(VEL_BB); w.dir(F); w.uzi(1); }
        return true;
    }
    if (id == 10) { // parallelPlatesBounceBackFreeSlip.cu: BB S, free slip N
        if (y == 0) { w.scheme(BOUNCE_BACK); w.dir(S); }
        else if (y == ny - 1) { w.scheme(FREE_SLIP); w.dir(N); }
        return true;
    }
    return false;
}

// distPoints2D casts only the first difference to float
// (src/CUDA/globalFunctions.h:158-162).
template <typename R>
R dist2d(R x1, R y1, R x2, R y2) {
    return (R)std::sqrt((float)(x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2));
}

// circularDuctInterpBounceBack.cu:35-121 / annularDuctInterpBounceBack.cu.
// Geometry evaluated in the reference's dfloat precision (double here).
// parallelPlatesRheometer.cu:35-136 (rheometer): the circular duct with the
// planes z = 0 and NZ-1 unused and gpuSchSpecial on z = 1 (BACK) and NZ-2
// (FRONT), set before the wall classification, which may override it.
template <typename R>
void duct_word(const or_desc& d, int x, int y, Word& w, bool annular, int z = -1) {
    const R Rout = (R)d.duct_outer_radius;
    const R xc = (R)(((d.nx - 1) / 2.0) + 0.5), yc = (R)(((d.ny - 1) / 2.0) + 0.5);
    const R xn = (R)(x + 0.5), yn = (R)(y + 0.5);
    const R dist = dist2d<R>(xn, yn, xc, yc);
    const int dirs[8][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}, {1, 1}, {-1, 1}, {-1, -1}, {1, -1}};
    if (!annular) {
        const bool rheometer = z >= 0;
        if (dist > Rout || (rheometer && (z == 0 || z == d.nz - 1))) { w.used(false); return; }
        if (rheometer) {
            if (z == 1) { w.scheme(SPECIAL); w.dir(B); }
            else if (z == d.nz - 2) { w.scheme(SPECIAL); w.dir(F); }
        }
        if (dist > (Rout - std::sqrt((float)2))) {
            w.scheme(INTERP_BB); w.dir(N);  // setIsInsideNodeInterpoBB(false) -> NORTH
        } else if (dist > (Rout - 2 * std::sqrt((float)2))) {
            w.save(true);                   // the unknown-population loop below runs for this ring too
        } else {
            return;
        }
        for (int i = 0; i < 8; i++) {
            const R xa = xn - dirs[i][0], ya = yn - dirs[i][1];
            if (xa < d.nx && xa > 0 && ya < d.ny && ya > 0) {
                if (dist2d<R>(xa, ya, xc, yc) > Rout) w.unknown_interp(i);
            } else {
                w.unknown_interp(i);
            }
        }
        return;
    }
    // annularDuctInterpBounceBack.cu: direction FRONT first, then the
    // classification; the unknown-population loop runs for every node.
    const R Rin = (R)d.duct_inner_radius;
    w.dir(F);
    if (dist > Rout || dist < Rin) {
        w.used(false);
    } else {
        if (dist > (Rout - std::sqrt((float)2))) { w.scheme(INTERP_BB); w.dir(N); }
        else if (dist > (Rout - 2 * std::sqrt((float)2))) { w.save(true); }
        else if (dist < Rin + std::sqrt((float)2)) { w.scheme(INTERP_BB); w.dir(S); }
        else if (dist < Rin + 2 * std::sqrt((float)2)) { w.save(true); }
    }
    for (int i = 0; i < 8; i++) {
        const R xa = xn - dirs[i][0], ya = yn - dirs[i][1];
        if (xa < d.nx && xa > 0 && ya < d.ny && ya > 0) {
            const R da = dist2d<R>(xa, ya, xc, yc);
            if (da > Rout || da < Rin) w.unknown_interp(i);
        } else {
            w.unknown_interp(i);
        }
    }
}

uint32_t build_word(const or_desc& d, int x, int y, int z) {
    Word w;
    w.used(true); w.save(false); w.scheme(BC_NULL); w.geom(0);
    w.uxi(0); w.uyi(0); w.uzi(0); w.rhoi(0);
    if (d.builder == 12) { duct_word<double>(d, x, y, w, false); return w.m; }
    if (d.builder == 13) { duct_word<double>(d, x, y, w, true); return w.m; }
    if (d.builder == 14) { duct_word<double>(d, x, y, w, false, z); return w.m; }
    if (short_chain_word(d.builder, x, y, z, d.nx, d.ny, d.nz, w)) return w.m;
    Act a[27];
    builder_actions(d.builder, a);
    const Act& act = a[classify(x, y, z, d.nx, d.ny, d.nz)];
    if (act.scheme >= 0) {
        w.scheme(act.scheme); w.dir(act.dir);
        if (act.uxi) w.uxi(act.uxi);
        if (act.uyi) w.uyi(act.uyi);
        if (act.uzi) w.uzi(act.uzi);
        if (act.rhoi) w.rhoi(act.rhoi);
    }
    return w.m;
}

// ---------------------------------------------------------------------------
// Bounce-back pairs per direction (dst <- src), D3Q19 part then D3Q27 extras
// (src/CUDA/boundaryConditionsSchemes/bounceBack.cu:26-499).
// Pinned by: tests derive the same sets from the rule "assign every incoming
// population whose opposite is outgoing" and compare.
// ---------------------------------------------------------------------------
struct BBList { int n19; int p19[7][2]; int n27; int p27[4][2]; };
const BBList BB_TABLE[26] = {
/*N  */ {5, {{4,3},{8,7},{12,11},{13,14},{18,17}}, 4, {{20,19},{22,21},{23,24},{26,25}}},
/*S  */ {5, {{3,4},{7,8},{11,12},{14,13},{17,18}}, 4, {{19,20},{21,22},{24,23},{25,26}}},
/*W  */ {5, {{1,2},{7,8},{9,10},{13,14},{15,16}}, 4, {{19,20},{21,22},{23,24},{26,25}}},
/*E  */ {5, {{2,1},{8,7},{10,9},{14,13},{16,15}}, 4, {{20,19},{22,21},{24,23},{25,26}}},
/*F  */ {5, {{6,5},{10,9},{12,11},{15,16},{17,18}}, 4, {{20,19},{21,22},{24,23},{26,25}}},
/*B  */ {5, {{5,6},{9,10},{11,12},{16,15},{18,17}}, 4, {{19,20},{22,21},{23,24},{25,26}}},
/*NW */ {7, {{1,2},{4,3},{9,10},{12,11},{13,14},{15,16},{18,17}}, 2, {{23,24},{26,25}}},
/*NE */ {7, {{2,1},{4,3},{8,7},{10,9},{12,11},{16,15},{18,17}}, 2, {{20,19},{22,21}}},
/*NF */ {7, {{4,3},{6,5},{8,7},{10,9},{12,11},{13,14},{15,16}}, 2, {{20,19},{26,25}}},
/*NB */ {7, {{4,3},{5,6},{8,7},{9,10},{13,14},{16,15},{18,17}}, 2, {{22,21},{23,24}}},
/*SW */ {7, {{1,2},{3,4},{7,8},{9,10},{11,12},{15,16},{17,18}}, 2, {{19,20},{21,22}}},
/*SE */ {7, {{2,1},{3,4},{10,9},{11,12},{14,13},{16,15},{17,18}}, 2, {{24,23},{25,26}}},
/*SF */ {7, {{3,4},{6,5},{7,8},{10,9},{14,13},{15,16},{17,18}}, 2, {{21,22},{24,23}}},
/*SB */ {7, {{3,4},{5,6},{7,8},{9,10},{11,12},{14,13},{16,15}}, 2, {{19,20},{25,26}}},
/*WF */ {7, {{1,2},{6,5},{7,8},{12,11},{13,14},{15,16},{17,18}}, 2, {{21,22},{26,25}}},
/*WB */ {7, {{1,2},{5,6},{7,8},{9,10},{11,12},{13,14},{18,17}}, 2, {{19,20},{23,24}}},
/*EF */ {7, {{2,1},{6,5},{8,7},{10,9},{12,11},{14,13},{17,18}}, 2, {{20,19},{24,23}}},
/*EB */ {7, {{2,1},{5,6},{8,7},{11,12},{14,13},{16,15},{18,17}}, 2, {{22,21},{25,26}}},
/*NWF*/ {6, {{1,2},{4,3},{6,5},{12,11},{13,14},{15,16}}, 1, {{26,25}}},
/*NWB*/ {6, {{1,2},{4,3},{5,6},{9,10},{13,14},{18,17}}, 1, {{23,24}}},
/*NEF*/ {6, {{2,1},{4,3},{6,5},{8,7},{10,9},{12,11}}, 1, {{20,19}}},
/*NEB*/ {6, {{2,1},{4,3},{5,6},{8,7},{16,15},{18,17}}, 1, {{22,21}}},
/*SWF*/ {6, {{1,2},{3,4},{6,5},{7,8},{15,16},{17,18}}, 1, {{21,22}}},
/*SWB*/ {6, {{1,2},{3,4},{5,6},{7,8},{9,10},{11,12}}, 1, {{19,20}}},
/*SEF*/ {6, {{2,1},{3,4},{6,5},{10,9},{14,13},{17,18}}, 1, {{24,23}}},
/*SEB*/ {6, {{2,1},{3,4},{5,6},{11,12},{14,13},{16,15}}, 1, {{25,26}}},
};

// Free-slip reflections (dst <- src at the node offset by (dx,dy,dz)),
// src/CUDA/boundaryConditionsSchemes/freeSlip.cu:26-168. z offsets wrap modulo NZ.
struct FSEntry { int dst, src, dx, dy, dz; };
const FSEntry FS_TABLE[6][9] = {
/*N*/ {{4,3,0,0,0},{8,14,1,0,0},{12,17,0,0,1},{13,7,-1,0,0},{18,11,0,0,-1},
       {20,24,1,0,1},{22,25,1,0,-1},{23,19,-1,0,-1},{26,21,-1,0,1}},
/*S*/ {{3,4,0,0,0},{7,13,-1,0,0},{11,18,0,0,-1},{14,8,1,0,0},{17,12,0,0,1},
       {19,23,-1,0,-1},{21,26,-1,0,1},{24,20,1,0,1},{25,22,1,0,-1}},
/*W*/ {{1,2,0,0,0},{7,14,0,-1,0},{9,16,0,0,-1},{13,8,0,1,0},{15,10,0,0,1},
       {19,25,0,-1,-1},{21,24,0,-1,1},{23,22,0,1,-1},{26,20,0,1,1}},
/*E*/ {{2,1,0,0,0},{8,13,0,1,0},{10,15,0,0,1},{14,7,0,-1,0},{16,9,0,0,-1},
       {20,26,0,1,1},{22,23,0,1,-1},{24,21,0,-1,1},{25,19,0,-1,-1}},
/*F*/ {{6,5,0,0,0},{10,16,1,0,0},{12,18,0,1,0},{15,9,-1,0,0},{17,11,0,-1,0},
       {20,22,1,1,0},{21,19,-1,-1,0},{24,25,1,-1,0},{26,23,-1,1,0}},
/*B*/ {{5,6,0,0,0},{9,15,-1,0,0},{11,17,0,-1,0},{16,10,1,0,0},{18,12,0,1,0},
       {19,21,-1,-1,0},{22,20,1,1,0},{23,26,-1,1,0},{25,24,1,-1,0}},
};

// ---------------------------------------------------------------------------
// Lattice state
// ---------------------------------------------------------------------------
template <typename R>
struct Lattice {
    or_desc d;
    int nx, ny, nz, q;
    size_t nnodes, nplane;
    std::vector<R> pop, popAux;
    std::vector<uint32_t> map;
    std::vector<size_t> bc_list;
    std::vector<R> rho, ux, uy, uz, omega;
    std::vector<R> visc_turb;   // LES_EXPORT_VISC_TURBULENT: macr.visc_turb
    std::vector<R> force_x, force_y, force_z;   // per-node force (IBM builds only)
    R w[27];

    size_t ip(int x, int y, int z, int i) const {
        return (size_t)nx * ((size_t)ny * ((size_t)(nz + 1) * i + z) + y) + x;
    }
    size_t is(int x, int y, int z) const { return (size_t)nx * ((size_t)ny * z + y) + x; }

    explicit Lattice(const or_desc& desc) : d(desc), nx(desc.nx), ny(desc.ny), nz(desc.nz), q(desc.q) {
        nnodes = (size_t)nx * ny * nz;
        nplane = (size_t)nx * ny * (nz + 1);
        pop.assign(nplane * q, R(0));
        popAux.assign(nplane * q, R(0));
        rho.assign(nnodes, R(0)); ux.assign(nnodes, R(0)); uy.assign(nnodes, R(0));
        uz.assign(nnodes, R(0)); omega.assign(nnodes, R(0));
        visc_turb.assign(nnodes, R(0));   // INIT_VISC_TURB (lbmInitialization.cu:114-115) is 0
        map.resize(nnodes);
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) map[is(x, y, z)] = build_word(d, x, y, z);
        // setupBoundaryConditionsInfo: used nodes with a BC scheme, z->y->x order
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    const uint32_t m = map[is(x, y, z)];
                    if (w_used(m) && w_scheme(m) != BC_NULL) bc_list.push_back(is(x, y, z));
                }
        if (q == 19) {
            const R W0 = (R)(1.0 / 3), W1 = (R)(1.0 / 18), W2 = (R)(1.0 / 36);
            w[0] = W0;
            for (int i = 1; i < 7; i++) w[i] = W1;
            for (int i = 7; i < 19; i++) w[i] = W2;
        } else {
            const R W0 = (R)(8.0 / 27), W1 = (R)(2.0 / 27), W2 = (R)(1.0 / 54), W3 = (R)(1.0 / 216);
            w[0] = W0;
            for (int i = 1; i < 7; i++) w[i] = W1;
            for (int i = 7; i < 19; i++) w[i] = W2;
            for (int i = 19; i < 27; i++) w[i] = W3;
        }
    }

    // gpu_f_eq (src/CUDA/globalFunctions.h:50-56) as used by gpuInitialization
    // (src/CUDA/lbmInitialization.cu:48-90).
    R feq(int i, R r, R vx, R vy, R vz) const {
        const R rhow = w[i] * r;
        const R uc3 = 3 * (vx * CX[i] + vy * CY[i] + vz * CZ[i]);
        const R p1_muu = 1 - 1.5 * (vx * vx + vy * vy + vz * vz);
        return (rhow * (p1_muu + uc3 * (1.0 + uc3 * 0.5)));
    }

    void init(const double* r0, const double* vx0, const double* vy0, const double* vz0) {
        for (int z = 0; z <= nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++) {
                    R r = 1, vx = 0, vy = 0, vz = 0;
                    if (z < nz) {
                        const size_t k = is(x, y, z);
                        r = r0 ? (R)r0[k] : (R)d.rho0;
                        vx = vx0 ? (R)vx0[k] : R(0);
                        vy = vy0 ? (R)vy0[k] : R(0);
                        vz = vz0 ? (R)vz0[k] : R(0);
                        rho[k] = r; ux[k] = vx; uy[k] = vy; uz[k] = vz;
                        omega[k] = 0;   // gpuMacrInitValue (src/CUDA/lbmInitialization.cu:110-112)
                    }
                    for (int i = 0; i < q; i++) {
                        const R fe = feq(i, r, vx, vy, vz);
                        pop[ip(x, y, z, i)] = fe;
                        popAux[ip(x, y, z, i)] = fe;
                    }
                }
    }

    // calcOmega (src/CUDA/NNF/nnf.h:32-87), plus the optional tau clamp.
    R calc_omega(R omegaOld, R stressMag) const {
        R om;
        if (d.rheo == 2) {
            R eta = ((1 / omegaOld) - 0.5) / 3.0;
            const R gammaDot = (1 - 0.5 * omegaOld) * stressMag / eta;
            if (gammaDot <= (R)d.gamma_0) eta = (R)d.k_consistency;
            else eta = (R)d.k_consistency * std::pow(gammaDot, (R)((R)d.n_index - 1));
            om = 1 / (0.5 + 3 * eta);
        } else {
            const R omega_p = (R)(1 / (3 * (R)d.eta_p + 0.5));
            const R t = (1 - (R)d.s_y / stressMag);
            om = omega_p * ((0.0 > t) ? 0.0 : t);
        }
        if (d.omega_clamp) {
            const R tau = 1 / om;
            if (!(tau >= (R)d.tau_min)) om = (R)(1 / (R)d.tau_min);
            else if (tau > (R)d.tau_max) om = (R)(1 / (R)d.tau_max);
        }
        return om;
    }

    // Sum of f over an index list, left to right.
    static R sum_list(const R* f, const int* l, int n) {
        R s = f[l[0]];
        for (int k = 1; k < n; k++) s += f[l[k]];
        return s;
    }

    // Per-node body of gpuMacrCollisionStream (src/CUDA/lbm.cu:3-504),
    // collision only: fin -> fout, macroscopics out. omega_io is the per-node
    // omega state (power law) and is updated in place.
    void collide(R* f, R fxVar, R fyVar, R fzVar, R* omega_io, R* m_out) const {
        const bool q27 = (q == 27);
        R rhoVar = f[0];
        for (int i = 1; i < q; i++) rhoVar += f[i];
        const R invRho = 1 / rhoVar;
        static const int XP19[] = {1, 7, 9, 13, 15}, XM19[] = {2, 8, 10, 14, 16};
        static const int YP19[] = {3, 7, 11, 14, 17}, YM19[] = {4, 8, 12, 13, 18};
        static const int ZP19[] = {5, 9, 11, 16, 18}, ZM19[] = {6, 10, 12, 15, 17};
        static const int XP27[] = {1, 7, 9, 13, 15, 19, 21, 23, 26}, XM27[] = {2, 8, 10, 14, 16, 20, 22, 24, 25};
        static const int YP27[] = {3, 7, 11, 14, 17, 19, 21, 24, 25}, YM27[] = {4, 8, 12, 13, 18, 20, 22, 23, 26};
        static const int ZP27[] = {5, 9, 11, 16, 18, 19, 22, 23, 25}, ZM27[] = {6, 10, 12, 15, 17, 20, 21, 24, 26};
        const int nl = q27 ? 9 : 5;
        const R uxVar = ((sum_list(f, q27 ? XP27 : XP19, nl)) - (sum_list(f, q27 ? XM27 : XM19, nl)) + 0.5 * fxVar) * invRho;
        const R uyVar = ((sum_list(f, q27 ? YP27 : YP19, nl)) - (sum_list(f, q27 ? YM27 : YM19, nl)) + 0.5 * fyVar) * invRho;
        const R uzVar = ((sum_list(f, q27 ? ZP27 : ZP19, nl)) - (sum_list(f, q27 ? ZM27 : ZM19, nl)) + 0.5 * fzVar) * invRho;

        const R fxVar_D3 = fxVar / 3, fyVar_D3 = fyVar / 3, fzVar_D3 = fzVar / 3;
        const R p1_muu15 = 1 - 1.5 * (uxVar * uxVar + uyVar * uyVar + uzVar * uzVar);
        const R W0 = w[0], W1 = w[1], W2 = w[7], W3 = q27 ? w[19] : R(0);
        const R rhoW0 = rhoVar * W0, rhoW1 = rhoVar * W1, rhoW2 = rhoVar * W2, rhoW3 = rhoVar * W3;
        const R W1t9d2 = W1 * 4.5, W2t9d2 = W2 * 4.5, W3t9d2 = W3 * 4.5;
        const R ux3 = 3 * uxVar, uy3 = 3 * uyVar, uz3 = 3 * uzVar;
        const R ux3ux3d2 = ux3 * ux3 * 0.5, ux3uy3 = ux3 * uy3, ux3uz3 = ux3 * uz3;
        const R uy3uy3d2 = uy3 * uy3 * 0.5, uy3uz3 = uy3 * uz3, uz3uz3d2 = uz3 * uz3 * 0.5;
        R terms[11];
        R mt;
        R aux;
        const R cs2 = 1.0 / 3.0;

        // second moments (src/CUDA/lbm.cu:149-190)
        static const int XX19[] = {1, 2, 7, 8, 9, 10, 13, 14, 15, 16};
        static const int YY19[] = {3, 4, 7, 8, 11, 12, 13, 14, 17, 18};
        static const int ZZ19[] = {5, 6, 9, 10, 11, 12, 15, 16, 17, 18};
        R sXX = sum_list(f, XX19, 10), sYY = sum_list(f, YY19, 10), sZZ = sum_list(f, ZZ19, 10);
        R sXY = f[7] + f[8] - f[13] - f[14];
        R sXZ = f[9] + f[10] - f[15] - f[16];
        R sYZ = f[11] + f[12] - f[17] - f[18];
        if (q27) {
            for (int i = 19; i < 27; i++) { sXX += f[i]; }
            for (int i = 19; i < 27; i++) { sYY += f[i]; }
            for (int i = 19; i < 27; i++) { sZZ += f[i]; }
            sXY = sXY + f[19] + f[20] + f[21] + f[22] - f[23] - f[24] - f[25] - f[26];
            sXZ = sXZ + f[19] + f[20] - f[21] - f[22] + f[23] + f[24] - f[25] - f[26];
            sYZ = sYZ + f[19] + f[20] - f[21] - f[22] - f[23] - f[24] + f[25] + f[26];
        }
        const R mXX = sXX - rhoVar * (uxVar * uxVar + cs2);
        const R mYY = sYY - rhoVar * (uyVar * uyVar + cs2);
        const R mZZ = sZZ - rhoVar * (uzVar * uzVar + cs2);
        const R mXYt2 = (sXY - rhoVar * uxVar * uyVar) * 2;
        const R mXZt2 = (sXZ - rhoVar * uxVar * uzVar) * 2;
        const R mYZt2 = (sYZ - rhoVar * uyVar * uzVar) * 2;

        R omegaVar = (R)(1.0 / (R)d.tau);
        if (d.rheo != 0 || d.les_const != 0) {
            // stress magnitude (src/CUDA/lbm.cu:192-208)
            const R uFxxd2 = uxVar * fxVar, uFyyd2 = uyVar * fyVar, uFzzd2 = uzVar * fzVar;
            const R uFxyd2 = (uxVar * fyVar + uyVar * fxVar) / 2;
            const R uFxzd2 = (uxVar * fzVar + uzVar * fxVar) / 2;
            const R uFyzd2 = (uyVar * fzVar + uzVar * fyVar) / 2;
            const R stressMag = std::sqrt(0.5 * (
                (mXX + uFxxd2) * (mXX + uFxxd2) + (mYY + uFyyd2) * (mYY + uFyyd2) +
                (mZZ + uFzzd2) * (mZZ + uFzzd2) +
                2 * ((mXYt2 / 2 + uFxyd2) * (mXYt2 / 2 + uFxyd2) +
                     (mXZt2 / 2 + uFxzd2) * (mXZt2 / 2 + uFxzd2) +
                     (mYZt2 / 2 + uFyzd2) * (mYZt2 / 2 + uFyzd2))));
            if (d.rheo == 1) {
                omegaVar = calc_omega((R)(1 / (3 * (R)d.eta_p + 0.5)), stressMag);
            } else if (d.rheo == 2) {
                omegaVar = calc_omega(*omega_io, stressMag);
                *omega_io = omegaVar;
            }
            if (d.les_const != 0) {
                // LES_MODEL (lbm.cu:226-242): implicit Smagorinsky eddy relaxation
                const R TAU = (R)d.tau;
                const R tau_t = 0.5 * std::sqrt(TAU * TAU + (R)d.les_const * stressMag) - 0.5 * TAU;
                m_out[5] = tau_t / 3.0;   // visc_turb_var (lbm.cu:229)
                omegaVar = 1.0 / (TAU + tau_t);
            }
        }
        const R t_omegaVar = 1 - omegaVar;
        const R tt_omegaVar = 1 - 0.5 * omegaVar;
        m_out[0] = rhoVar; m_out[1] = uxVar; m_out[2] = uyVar; m_out[3] = uzVar; m_out[4] = omegaVar;

        const bool bgk = d.collision == 1;
        if (bgk) {
            // BGK: fNode *= T_OMEGA, then += OMEGA rho w terms
            // (recursiveCollisionForceNoReg.txt:38-104)
            for (int i = 0; i < q; i++) f[i] *= t_omegaVar;
        } else {
        // regularized non-equilibrium (src/CUDA/lbm.cu:263-325)
        terms[0] = -mXX / 3 - mYY / 3 - mZZ / 3;
        terms[1] = terms[0] + (-fxVar_D3 + mXX);
        terms[2] = terms[0] + (fxVar_D3 + mXX);
        terms[3] = terms[0] + (-fyVar_D3 + mYY);
        terms[4] = terms[0] + (fyVar_D3 + mYY);
        if (q27) {
            terms[5] = terms[1] + (-fyVar_D3 + mXYt2 + mYY);
            terms[6] = terms[2] + (fyVar_D3 + mXYt2 + mYY);
            terms[7] = terms[1] + (-fzVar_D3 + mXZt2 + mZZ);
            terms[8] = terms[2] + (fzVar_D3 + mXZt2 + mZZ);
            terms[9] = terms[3] + (-fzVar_D3 + mYZt2 + mZZ);
            terms[10] = terms[4] + (fzVar_D3 + mYZt2 + mZZ);
        }
        mt = W0 * 4.5;
        f[0] = mt * terms[0];
        mt = W1t9d2;
        f[1] = mt * terms[1]; f[2] = mt * terms[2]; f[3] = mt * terms[3]; f[4] = mt * terms[4];
        f[5] = mt * (terms[0] + (-fzVar_D3 + mZZ));
        f[6] = mt * (terms[0] + (fzVar_D3 + mZZ));
        mt = W2t9d2;
        f[7] = mt * (terms[1] + (-fyVar_D3 + mXYt2 + mYY));
        f[8] = mt * (terms[2] + (fyVar_D3 + mXYt2 + mYY));
        f[9] = mt * (terms[1] + (-fzVar_D3 + mXZt2 + mZZ));
        f[10] = mt * (terms[2] + (fzVar_D3 + mXZt2 + mZZ));
        f[11] = mt * (terms[3] + (-fzVar_D3 + mYZt2 + mZZ));
        f[12] = mt * (terms[4] + (fzVar_D3 + mYZt2 + mZZ));
        f[13] = mt * (terms[1] + (fyVar_D3 - mXYt2 + mYY));
        f[14] = mt * (terms[2] + (-fyVar_D3 - mXYt2 + mYY));
        f[15] = mt * (terms[1] + (fzVar_D3 - mXZt2 + mZZ));
        f[16] = mt * (terms[2] + (-fzVar_D3 - mXZt2 + mZZ));
        f[17] = mt * (terms[3] + (fzVar_D3 - mYZt2 + mZZ));
        f[18] = mt * (terms[4] + (-fzVar_D3 - mYZt2 + mZZ));
        if (q27) {
            mt = W3t9d2;
            f[19] = mt * (terms[5] + (-fzVar_D3 + mXZt2 + mYZt2 + mZZ));
            f[20] = mt * (terms[6] + (fzVar_D3 + mXZt2 + mYZt2 + mZZ));
            f[21] = mt * (terms[5] + (fzVar_D3 - mXZt2 - mYZt2 + mZZ));
            f[22] = mt * (terms[6] + (-fzVar_D3 - mXZt2 - mYZt2 + mZZ));
            f[23] = mt * (terms[7] + (fyVar_D3 - mXYt2 + mYY - mYZt2));
            f[24] = mt * (terms[8] + (-fyVar_D3 - mXYt2 + mYY - mYZt2));
            f[25] = mt * (terms[9] + (fxVar_D3 + mXX - mXYt2 - mXZt2));
            f[26] = mt * (terms[10] + (-fxVar_D3 + mXX - mXYt2 - mXZt2));
        }
        for (int i = 0; i < q; i++) f[i] *= t_omegaVar;
        }

        // equilibrium (src/CUDA/lbm.cu:341-401); BGK scales it by omega
        const R eqs = bgk ? omegaVar : R(1);
        terms[0] = p1_muu15;
        terms[1] = terms[0] + (ux3 + ux3ux3d2);
        terms[2] = terms[0] + (-ux3 + ux3ux3d2);
        terms[3] = terms[0] + (uy3 + uy3uy3d2);
        terms[4] = terms[0] + (-uy3 + uy3uy3d2);
        if (q27) {
            terms[5] = terms[1] + (uy3 + ux3uy3 + uy3uy3d2);
            terms[6] = terms[2] + (-uy3 + ux3uy3 + uy3uy3d2);
            terms[7] = terms[1] + (uz3 + ux3uz3 + uz3uz3d2);
            terms[8] = terms[2] + (-uz3 + ux3uz3 + uz3uz3d2);
            terms[9] = terms[3] + (uz3 + uy3uz3 + uz3uz3d2);
            terms[10] = terms[4] + (-uz3 + uy3uz3 + uz3uz3d2);
        }
        mt = eqs * rhoW0;
        f[0] += mt * terms[0];
        mt = eqs * rhoW1;
        f[1] += mt * terms[1]; f[2] += mt * terms[2]; f[3] += mt * terms[3]; f[4] += mt * terms[4];
        f[5] += mt * (terms[0] + (uz3 + uz3uz3d2));
        f[6] += mt * (terms[0] + (-uz3 + uz3uz3d2));
        mt = eqs * rhoW2;
        f[7] += mt * (terms[1] + (uy3 + ux3uy3 + uy3uy3d2));
        f[8] += mt * (terms[2] + (-uy3 + ux3uy3 + uy3uy3d2));
        f[9] += mt * (terms[1] + (uz3 + ux3uz3 + uz3uz3d2));
        f[10] += mt * (terms[2] + (-uz3 + ux3uz3 + uz3uz3d2));
        f[11] += mt * (terms[3] + (uz3 + uy3uz3 + uz3uz3d2));
        f[12] += mt * (terms[4] + (-uz3 + uy3uz3 + uz3uz3d2));
        f[13] += mt * (terms[1] + (-uy3 - ux3uy3 + uy3uy3d2));
        f[14] += mt * (terms[2] + (uy3 - ux3uy3 + uy3uy3d2));
        f[15] += mt * (terms[1] + (-uz3 - ux3uz3 + uz3uz3d2));
        f[16] += mt * (terms[2] + (uz3 - ux3uz3 + uz3uz3d2));
        f[17] += mt * (terms[3] + (-uz3 - uy3uz3 + uz3uz3d2));
        f[18] += mt * (terms[4] + (uz3 - uy3uz3 + uz3uz3d2));
        if (q27) {
            mt = eqs * rhoW3;
            f[19] += mt * (terms[5] + (uz3 + ux3uz3 + uy3uz3 + uz3uz3d2));
            f[20] += mt * (terms[6] + (-uz3 + ux3uz3 + uy3uz3 + uz3uz3d2));
            f[21] += mt * (terms[5] + (-uz3 - ux3uz3 - uy3uz3 + uz3uz3d2));
            f[22] += mt * (terms[6] + (uz3 - ux3uz3 - uy3uz3 + uz3uz3d2));
            f[23] += mt * (terms[7] + (-uy3 - ux3uy3 + uy3uy3d2 - uy3uz3));
            f[24] += mt * (terms[8] + (uy3 - ux3uy3 + uy3uy3d2 - uy3uz3));
            f[25] += mt * (terms[9] + (-ux3 + ux3ux3d2 - ux3uy3 - ux3uz3));
            f[26] += mt * (terms[10] + (ux3 + ux3ux3d2 - ux3uy3 - ux3uz3));
        }

        // Guo force (src/CUDA/lbm.cu:403-461)
        terms[0] = -fxVar * ux3 - fyVar * uy3 - fzVar * uz3;
        terms[1] = terms[0] + (fxVar * (3 * ux3 + 3));
        terms[2] = terms[0] + (fyVar * (3 * uy3 + 3));
        terms[3] = terms[1] + (fxVar * (3 * uy3) + fyVar * (3 * ux3 + 3 * uy3 + 3));
        terms[4] = terms[1] + (fxVar * (3 * uz3) + fzVar * (3 * ux3 + 3 * uz3 + 3));
        terms[5] = terms[2] + (fyVar * (3 * uz3) + fzVar * (3 * uy3 + 3 * uz3 + 3));
        if (q27)
            terms[6] = terms[3] + (fxVar * (3 * uz3) + fyVar * (3 * uz3) + fzVar * (3 * ux3 + 3 * uy3 + 3 * uz3 + 3));
        mt = W0 * tt_omegaVar;
        f[0] += mt * terms[0];
        mt = W1 * tt_omegaVar;
        f[1] += mt * terms[1];
        f[2] += mt * (terms[1] + (fxVar * (-6)));
        f[3] += mt * terms[2];
        f[4] += mt * (terms[2] + (fyVar * (-6)));
        aux = terms[0] + (fzVar * (3 * uz3 + 3));
        f[5] += mt * aux;
        f[6] += mt * (aux + (fzVar * (-6)));
        mt = W2 * tt_omegaVar;
        f[7] += mt * terms[3];
        f[8] += mt * (terms[3] + (fxVar * (-6) + fyVar * (-6)));
        f[9] += mt * terms[4];
        f[10] += mt * (terms[4] + (fxVar * (-6) + fzVar * (-6)));
        f[11] += mt * (terms[5]);
        f[12] += mt * (terms[5] + (fyVar * (-6) + fzVar * (-6)));
        aux = terms[3] + (fxVar * (-6 * uy3) + fyVar * (-6 * ux3 - 6));
        f[13] += mt * (aux);
        f[14] += mt * (aux + (fxVar * (-6) + fyVar * (6)));
        aux = terms[4] + (fxVar * (-6 * uz3) + fzVar * (-6 * ux3 - 6));
        f[15] += mt * aux;
        f[16] += mt * (aux + (fxVar * (-6) + fzVar * (6)));
        aux = terms[5] + (fyVar * (-6 * uz3) + fzVar * (-6 * uy3 - 6));
        f[17] += mt * aux;
        f[18] += mt * (aux + (fyVar * (-6) + fzVar * (6)));
        if (q27) {
            mt = W3 * tt_omegaVar;
            f[19] += mt * terms[6];
            f[20] += mt * (terms[6] + (fxVar * (-6) + fyVar * (-6) + fzVar * (-6)));
            aux = terms[6] + (fxVar * (-6 * uz3) + fyVar * (-6 * uz3) + fzVar * (-6 * ux3 - 6 * uy3 - 6));
            f[21] += mt * aux;
            f[22] += mt * (aux + (fxVar * (-6) + fyVar * (-6) + fzVar * (6)));
            aux = terms[6] + (fxVar * (-6 * uy3) + fyVar * (-6 * ux3 - 6 * uz3 - 6) + fzVar * (-6 * uy3));
            f[23] += mt * aux;
            f[24] += mt * (aux + (fxVar * (-6) + fyVar * (6) + fzVar * (-6)));
            aux = terms[6] + (fxVar * (-6 * uy3 - 6 * uz3 - 6) + fyVar * (-6 * ux3) + fzVar * (-6 * ux3));
            f[25] += mt * aux;
            f[26] += mt * (aux + (fxVar * (6) + fyVar * (-6) + fzVar * (-6)));
        }
    }

    // ---- Boundary schemes (read popAux = post-stream, pop = post-collision) ----
    R& PS(int x, int y, int z, int i) { return popAux[ip(x, y, z, i)]; }
    R PC(int x, int y, int z, int i) const { return pop[ip(x, y, z, i)]; }

    void bc_bounce_back(int dir, int geom, int x, int y, int z) {
        if (dir >= NW && geom != 0) return;   // convex edges/corners are skipped
        const BBList& l = BB_TABLE[dir];
        for (int k = 0; k < l.n19; k++) PS(x, y, z, l.p19[k][0]) = PC(x, y, z, l.p19[k][1]);
        if (q == 27)
            for (int k = 0; k < l.n27; k++) PS(x, y, z, l.p27[k][0]) = PC(x, y, z, l.p27[k][1]);
    }

    void bc_free_slip(int dir, int x, int y, int z) {
        if (dir > B) return;
        const int n = (q == 27) ? 9 : 5;
        for (int k = 0; k < n; k++) {
            const FSEntry& e = FS_TABLE[dir][k];
            const int xs = (x + e.dx + nx) % nx, ys = (y + e.dy + ny) % ny, zs = (z + e.dz + nz) % nz;
            PS(x, y, z, e.dst) = PC(xs, ys, zs, e.src);
        }
    }

    // D3Q19_VelZouHe.cu:27-172, including its non-textbook lines.
    void bc_vel_zouhe(int dir, int x, int y, int z, R ux_w, R uy_w, R uz_w) {
        if (q != 19) return;
        auto f = [&](int i) -> R& { return PS(x, y, z, i); };
        switch (dir) {
        case N: {
            const R rho_w = (1 / (1 + uy_w)) * (f(1) + f(2) + f(5) + f(6) + f(9) + f(15) + f(16) + f(10)
                + f(0) + 2 * (f(3) + f(7) + f(14) + f(11) + f(17)));
            const R nyx = 0.5 * (f(1) + f(9) + f(15) - (f(2) + f(16) + f(10))) - ux_w * rho_w / 3;
            const R nyz = 0.5 * (f(5) + f(9) + f(16) - (f(6) + f(15) + f(10))) - uz_w * rho_w / 3;
            f(4) = f(3) + rho_w * (-uy_w) / 3;
            f(8) = f(7) + rho_w * (-uy_w - ux_w) / 6 + nyx;
            f(13) = f(14) + rho_w * (-uy_w + ux_w) / 6 - nyx;
            f(12) = f(11) + rho_w * (-uy_w - uz_w) / 6 + nyz;
            f(18) = f(17) + rho_w * (-uy_w + uz_w) / 6 - nyz;
            break; }
        case S: {
            const R rho_w = (1 / (1 - uy_w)) * (f(1) + f(2) + f(5) + f(6) + f(9) + f(15) + f(16) + f(10)
                + f(0) + 2 * (f(4) + f(13) + f(8) + f(18) + f(12)));
            const R nyx = 0.5 * (f(1) + f(9) + f(15) - (f(2) + f(16) + f(10))) - ux_w * rho_w / 3;
            const R nyz = 0.5 * (f(5) + f(9) + f(16) - (f(6) + f(15) + f(10))) - uz_w * rho_w / 3;
            f(3) = f(4) + rho_w * uy_w / 3;
            f(7) = f(8) + rho_w * (uy_w + ux_w) / 6 - nyx;
            f(14) = f(13) + rho_w * (uy_w - ux_w) / 6 + nyx;
            f(11) = f(12) + rho_w * (uy_w + uz_w) / 6 - nyz;
            f(17) = f(18) + rho_w * (uy_w - uz_w) / 6 + nyz;
            break; }
        case W: {
            const R rho_w = (1 / (1 - ux_w)) * (f(3) + f(4) + f(5) + f(6) + f(11) + f(17) + f(18) + f(12)
                + f(0) + 2 * (f(2) + f(14) + f(8) + f(16) + f(10)));
            const R nxy = 0.5 * (f(3) + f(11) + f(17) - (f(4) + f(18) + f(12))) - uy_w * rho_w / 3;
            const R nxz = 0.5 * (f(5) + f(18) + f(11) - (f(6) + f(17) + f(12))) - uz_w * rho_w / 3;
            f(1) = f(2) + rho_w * ux_w / 3;
            f(13) = f(14) + rho_w * (ux_w - uy_w) / 6 + nxy;
            f(7) = f(8) + rho_w * (ux_w + uy_w) / 6 - nxy;
            f(9) = f(10) + rho_w * (ux_w + uz_w) / 6 - nxz;
            f(15) = f(16) + rho_w * (ux_w - uy_w) / 6 + nxz;   // reference uses uy_w here
            break; }
        case E: {
            const R rho_w = (1 / (1 + ux_w)) * (f(3) + f(4) + f(5) + f(6) + f(11) + f(17) + f(18) + f(12)
                + f(0) + 2 * (f(2) + f(14) + f(8) + f(16) + f(10)));
            const R nxy = 0.5 * (f(3) + f(11) + f(17) - (f(4) + f(18) + f(12))) - uy_w * rho_w / 3;
            const R nxz = 0.5 * (f(5) + f(18) + f(11) - (f(6) + f(17) + f(12))) - uz_w * rho_w / 3;
            // reference builds the unknowns from themselves
            f(2) = f(2) + rho_w * ux_w / 3;
            f(14) = f(14) + rho_w * (ux_w - uy_w) / 6 + nxy;
            f(8) = f(8) + rho_w * (ux_w + uy_w) / 6 - nxy;
            f(10) = f(10) + rho_w * (ux_w + uz_w) / 6 - nxz;
            f(16) = f(16) + rho_w * (ux_w - uy_w) / 6 + nxz;
            break; }
        case F: {
            const R rho_w = (1 / (1 + uz_w)) * (f(1) + f(2) + f(3) + f(4) + f(7) + f(14) + f(8) + f(13)
                + f(0) + 2 * (f(5) + f(9) + f(16) + f(11) + f(18)));
            const R nzx = 0.5 * (f(1) + f(7) + f(13) - (f(2) + f(14) + f(8))) - rho_w * ux_w / 3;
            const R nzy = 0.5 * (f(3) + f(7) + f(14) - (f(4) + f(13) + f(8))) - rho_w * uy_w / 3;
            f(6) = f(5) + rho_w * (-uz_w) / 3;
            f(15) = f(16) + rho_w * (-uz_w + ux_w) / 6 - nzx;
            f(10) = f(9) + rho_w * (-uz_w - ux_w) / 6 + nzx;
            f(17) = f(18) + rho_w * (-uz_w + uy_w) / 6 - nzy;
            f(12) = f(11) + rho_w * (-uz_w - uy_w) / 6 + nzy;
            break; }
        case B: {
            // reference: 1/(1+uz_w) (textbook 1/(1-uz_w)); f11/f18 use ux_w
            const R rho_w = (1 / (1 + uz_w)) * (f(1) + f(2) + f(3) + f(4) + f(7) + f(14) + f(8) + f(13)
                + f(0) + 2 * (f(6) + f(15) + f(10) + f(17) + f(12)));
            const R nzx = 0.5 * (f(1) + f(7) + f(13) - (f(2) + f(14) + f(8))) - rho_w * ux_w / 3;
            const R nzy = 0.5 * (f(3) + f(7) + f(14) - (f(4) + f(13) + f(8))) - rho_w * uy_w / 3;
            f(5) = f(6) + rho_w * (uz_w) / 3;
            f(9) = f(10) + rho_w * (uz_w + ux_w) / 6 - nzx;
            f(16) = f(15) + rho_w * (uz_w - ux_w) / 6 + nzx;
            f(11) = f(12) + rho_w * (uz_w + ux_w) / 6 - nzy;
            f(18) = f(17) + rho_w * (uz_w - ux_w) / 6 + nzy;
            break; }
        default: break;
        }
    }

    // D3Q19_PresZouHe.cu:27-176.
    void bc_pres_zouhe(int dir, int x, int y, int z, R rho_w) {
        if (q != 19) return;
        auto f = [&](int i) -> R& { return PS(x, y, z, i); };
        switch (dir) {
        case N: {
            const R uy_w = -1 + (1 / rho_w) * (f(1) + f(2) + f(5) + f(6) + f(9) + f(15) + f(16) + f(10)
                + f(0) + 2 * (f(3) + f(7) + f(14) + f(11) + f(17)));
            const R nyx = 0.5 * (f(1) + f(9) + f(15) - (f(2) + f(16) + f(10)));
            const R nyz = 0.5 * (f(5) + f(9) + f(16) - (f(6) + f(15) + f(10)));
            f(4) = f(3) + rho_w * (-uy_w) / 3;
            f(8) = f(7) + rho_w * (-uy_w) / 6 + nyx;
            f(13) = f(14) + rho_w * (-uy_w) / 6 - nyx;
            f(12) = f(11) + rho_w * (-uy_w) / 6 + nyz;
            f(18) = f(17) + rho_w * (-uy_w) / 6 - nyz;
            break; }
        case S: {
            const R uy_w = 1 - (1 / rho_w) * (f(1) + f(2) + f(5) + f(6) + f(9) + f(15) + f(16) + f(10)
                + f(0) + 2 * (f(4) + f(13) + f(8) + f(18) + f(12)));
            const R nyx = 0.5 * (f(1) + f(9) + f(15) - (f(2) + f(16) + f(10)));
            const R nyz = 0.5 * (f(5) + f(9) + f(16) - (f(6) + f(15) + f(10)));
            f(3) = f(4) + rho_w * uy_w / 3;
            f(7) = f(8) + rho_w * (uy_w) / 6 - nyx;
            f(14) = f(13) + rho_w * (uy_w) / 6 + nyx;
            f(11) = f(12) + rho_w * (uy_w) / 6 - nyz;
            f(17) = f(18) + rho_w * (uy_w) / 6 + nyz;
            break; }
        case W: {
            const R ux_w = 1 - (1 / rho_w) * (f(3) + f(4) + f(5) + f(6) + f(11) + f(17) + f(18) + f(12)
                + f(0) + 2 * (f(2) + f(14) + f(8) + f(16) + f(10)));
            const R nxy = 0.5 * (f(3) + f(11) + f(17) - (f(4) + f(18) + f(12)));
            const R nxz = 0.5 * (f(5) + f(18) + f(11) - (f(6) + f(17) + f(12)));
            f(1) = f(2) + rho_w * ux_w / 3;
            f(13) = f(14) + rho_w * (ux_w) / 6 + nxy;
            f(7) = f(8) + rho_w * (ux_w) / 6 - nxy;
            f(9) = f(10) + rho_w * (ux_w) / 6 - nxz;
            f(15) = f(16) + rho_w * (ux_w) / 6 + nxz;
            break; }
        case E: {
            const R ux_w = -1 + (1 / rho_w) * (f(3) + f(4) + f(5) + f(6) + f(11) + f(17) + f(18) + f(12)
                + f(0) + 2 * (f(2) + f(14) + f(8) + f(16) + f(10)));
            const R nxy = 0.5 * (f(3) + f(11) + f(17) - (f(4) + f(18) + f(12)));
            const R nxz = 0.5 * (f(5) + f(18) + f(11) - (f(6) + f(17) + f(12)));
            f(2) = f(2) + rho_w * ux_w / 3;
            f(14) = f(14) + rho_w * (ux_w) / 6 + nxy;
            f(8) = f(8) + rho_w * (ux_w) / 6 - nxy;
            f(10) = f(10) + rho_w * (ux_w) / 6 - nxz;
            f(16) = f(16) + rho_w * (ux_w) / 6 + nxz;
            break; }
        case F: {
            const R uz_w = -1 + (1 / rho_w) * (f(1) + f(2) + f(3) + f(4) + f(7) + f(14) + f(8) + f(13)
                + f(0) + 2 * (f(5) + f(9) + f(16) + f(11) + f(18)));
            const R nzx = 0.5 * (f(1) + f(7) + f(13) - (f(2) + f(14) + f(8)));
            const R nzy = 0.5 * (f(3) + f(7) + f(14) - (f(4) + f(13) + f(8)));
            f(6) = f(5) + rho_w * (-uz_w) / 3;
            f(15) = f(16) + rho_w * (-uz_w) / 6 - nzx;
            f(10) = f(9) + rho_w * (-uz_w) / 6 + nzx;
            f(17) = f(18) + rho_w * (-uz_w) / 6 - nzy;
            f(12) = f(11) + rho_w * (-uz_w) / 6 + nzy;
            break; }
        case B: {
            const R uz_w = 1 - (1 / rho_w) * (f(1) + f(2) + f(3) + f(4) + f(7) + f(14) + f(8) + f(13)
                + f(0) + 2 * (f(6) + f(10) + f(12) + f(15) + f(17)));
            const R nzx = 0.5 * (f(1) + f(7) + f(13) - (f(2) + f(14) + f(8)));
            const R nzy = 0.5 * (f(3) + f(7) + f(14) - (f(4) + f(13) + f(8)));
            f(5) = f(6) + rho_w * uz_w / 3;
            f(9) = f(10) + rho_w * uz_w / 6 - nzx;
            f(16) = f(15) + rho_w * uz_w / 6 + nzx;
            f(11) = f(12) + rho_w * uz_w / 6 - nzy;
            f(18) = f(17) + rho_w * uz_w / 6 + nzy;
            break; }
        default: break;
        }
    }

    // D3Q19_VelBounceBack.cu:27-149: f_in = f_out - 6 rho w (c_out . u_w), rho from post-stream pops.
    void bc_vel_bb(int dir, int x, int y, int z, R ux_w, R uy_w, R uz_w) {
        if (q != 19 || dir > B) return;
        auto f = [&](int i) -> R& { return PS(x, y, z, i); };
        R rho_w = f(0);
        for (int i = 1; i < 19; i++) rho_w += f(i);
        const R W1 = w[1], W2 = w[7];
        switch (dir) {
        case N:
            f(4) = f(3) - 6 * rho_w * W1 * (uy_w);
            f(8) = f(7) - 6 * rho_w * W2 * (uy_w + ux_w);
            f(12) = f(11) - 6 * rho_w * W2 * (uy_w + uz_w);
            f(13) = f(14) - 6 * rho_w * W2 * (uy_w - ux_w);
            f(18) = f(17) - 6 * rho_w * W2 * (uy_w - uz_w);
            break;
        case S:
            f(3) = f(4) - 6 * rho_w * W1 * (-uy_w);
            f(7) = f(8) - 6 * rho_w * W2 * (-uy_w - ux_w);
            f(11) = f(12) - 6 * rho_w * W2 * (-uy_w - uz_w);
            f(14) = f(13) - 6 * rho_w * W2 * (-uy_w + ux_w);
            f(17) = f(18) - 6 * rho_w * W2 * (-uy_w + uz_w);
            break;
        case W:
            f(1) = f(2) - 6 * rho_w * W1 * (-ux_w);
            f(7) = f(8) - 6 * rho_w * W2 * (-ux_w - uy_w);
            f(9) = f(10) - 6 * rho_w * W2 * (-ux_w - uz_w);
            f(13) = f(14) - 6 * rho_w * W2 * (-ux_w + uy_w);
            f(15) = f(16) - 6 * rho_w * W2 * (-ux_w + uz_w);
            break;
        case E:
            f(2) = f(1) - 6 * rho_w * W1 * (ux_w);
            f(8) = f(7) - 6 * rho_w * W2 * (ux_w + uy_w);
            f(10) = f(9) - 6 * rho_w * W2 * (ux_w + uz_w);
            f(14) = f(13) - 6 * rho_w * W2 * (ux_w - uy_w);
            f(16) = f(15) - 6 * rho_w * W2 * (ux_w - uz_w);
            break;
        case F:
            f(6) = f(5) - 6 * rho_w * W1 * (uz_w);
            f(10) = f(9) - 6 * rho_w * W2 * (uz_w + ux_w);
            f(12) = f(11) - 6 * rho_w * W2 * (uz_w + uy_w);
            f(15) = f(16) - 6 * rho_w * W2 * (uz_w - ux_w);
            f(17) = f(18) - 6 * rho_w * W2 * (uz_w - uy_w);
            break;
        case B:
            f(5) = f(6) - 6 * rho_w * W1 * (-uz_w);
            f(9) = f(10) - 6 * rho_w * W2 * (-uz_w - ux_w);
            f(11) = f(12) - 6 * rho_w * W2 * (-uz_w - uy_w);
            f(16) = f(15) - 6 * rho_w * W2 * (-uz_w + ux_w);
            f(18) = f(17) - 6 * rho_w * W2 * (-uz_w + uy_w);
            break;
        default: break;
        }
    }

    // ---- Interpolated bounce-back (src/CUDA/boundaryConditionsSchemes/interpolatedBounceBack.cu) ----
    static R q_a0(R x1, R y1, R x2, R y2, R Rr) {
        R y = y2, x = std::sqrt(Rr * Rr - y * y);
        if (x2 < 0) x = -x;
        return dist2d<R>(x, y, x1, y1);
    }
    static R q_ainf(R x1, R y1, R x2, R y2, R Rr) {
        R x = x2, y = std::sqrt(Rr * Rr - x * x);
        if (y2 < 0) y = -y;
        return dist2d<R>(x, y, x1, y1);
    }
    static R q_ap1(R x1, R y1, R x2, R y2, R Rr) {
        const R b = y1 - x1, dr = std::sqrt(2 * Rr * Rr - b * b);
        R x = (-b + dr) * 0.5;
        if ((x1 > x2 && (x > x1 || x < x2)) || (x1 < x2 && (x < x1 || x > x2))) x = (-b - dr) * 0.5;
        const R y = x + b;
        return dist2d<R>(x, y, x1, y1) / (R)1.41421356237309504880168872420969807856967187537;
    }
    static R q_am1(R x1, R y1, R x2, R y2, R Rr) {
        const R b = y1 + x1, dr = std::sqrt(2 * Rr * Rr - b * b);
        R x = (b + dr) * 0.5;
        if ((x1 > x2 && (x > x1 || x < x2)) || (x1 < x2 && (x < x1 || x > x2))) x = (b - dr) * 0.5;
        const R y = -x + b;
        return dist2d<R>(x, y, x1, y1) / (R)1.41421356237309504880168872420969807856967187537;
    }

    void bc_interp_bb(unsigned unknown, bool inside, int x, int y, int z) {
        const int zp1 = (z + 1) % nz, zm1 = (nz + z - 1) % nz;
        const R W1 = w[1], W2 = w[7], W3 = (q == 27) ? w[19] : R(0);
        const R xN = x - (nx - 1) / 2.0, yN = y - (ny - 1) / 2.0;
        const R rr = std::sqrt(xN * xN + yN * yN);
        const R c = xN / rr, s = yN / rr;
        R radius, uxw, uyw, uzw;
        const R wo = (R)(d.outer_rotation * ((R)z / (R)nz));
        if (inside) {
            radius = (R)d.duct_inner_radius;
            uxw = -(R)d.inner_rotation * radius * s; uyw = (R)d.inner_rotation * radius * c;
            uzw = (R)d.inner_velocity;
        } else {
            radius = (R)d.duct_outer_radius;
            uxw = -wo * radius * s; uyw = wo * radius * c; uzw = (R)d.outer_velocity;
        }
        const R as2 = 3.0;
        auto lowq = [&](R fb, R fa, R qq, R df) -> R { return (2 * qq * fb + (1 - 2 * qq) * fa + 2 * as2 * df); };
        auto highq = [&](R fi, R fj, R qq, R df) -> R { return ((0.5 / qq) * fi + ((qq - 0.5) / (qq)) * fj + as2 * df / qq); };
        // (d2q9 bit, dx, dy, q-function, axis pop, diag pops with their dz and weights)
        struct Grp { int bit, dx, dy, kind; int pops[3][3]; int n27; int pops27[2][3]; };
        // pops[k] = {dst, src(opp), dz}; velocity projections are evaluated below.
        const Grp G[8] = {
            {0, 1, 0, 0, {{1, 2, 0}, {9, 10, 1}, {15, 16, -1}}, 0, {{0,0,0},{0,0,0}}},
            {1, 0, 1, 1, {{3, 4, 0}, {11, 12, 1}, {17, 18, -1}}, 0, {{0,0,0},{0,0,0}}},
            {2, -1, 0, 0, {{2, 1, 0}, {16, 15, 1}, {10, 9, -1}}, 0, {{0,0,0},{0,0,0}}},
            {3, 0, -1, 1, {{4, 3, 0}, {18, 17, 1}, {12, 11, -1}}, 0, {{0,0,0},{0,0,0}}},
            {4, 1, 1, 2, {{7, 8, 0}, {0, 0, 0}, {0, 0, 0}}, 2, {{19, 20, 1}, {21, 22, -1}}},
            {5, -1, 1, 3, {{14, 13, 0}, {0, 0, 0}, {0, 0, 0}}, 2, {{25, 26, 1}, {24, 23, -1}}},
            {6, -1, -1, 2, {{8, 7, 0}, {0, 0, 0}, {0, 0, 0}}, 2, {{22, 21, 1}, {20, 19, -1}}},
            {7, 1, -1, 3, {{13, 14, 0}, {0, 0, 0}, {0, 0, 0}}, 2, {{23, 24, 1}, {26, 25, -1}}},
        };
        for (int g = 0; g < 8; g++) {
            const Grp& gr = G[g];
            if (!(unknown & (1u << gr.bit))) continue;
            R qq;
            const R x2 = xN - gr.dx, y2 = yN - gr.dy;
            if (gr.kind == 0) qq = q_a0(xN, yN, x2, y2, radius);
            else if (gr.kind == 1) qq = q_ainf(xN, yN, x2, y2, radius);
            else if (gr.kind == 2) qq = q_ap1(xN, yN, x2, y2, radius);
            else qq = q_am1(xN, yN, x2, y2, radius);
            const int np = (gr.dx != 0 && gr.dy != 0) ? 1 : 3;
            auto apply = [&](int dst, int src, int dz, R wgt) {
                const R cu = gr.dx * uxw + gr.dy * uyw + dz * uzw;
                const R df = wgt * cu;
                if (qq > 0.5) {
                    PS(x, y, z, dst) = highq(PC(x, y, z, src), PC(x, y, z, dst), qq, df);
                } else {
                    const int xa = x + gr.dx, ya = y + gr.dy;
                    const int za = (dz == 1) ? zp1 : ((dz == -1) ? zm1 : z);
                    PS(x, y, z, dst) = lowq(PC(x, y, z, src), PC(xa, ya, za, src), qq, df);
                }
            };
            for (int k = 0; k < np; k++) apply(gr.pops[k][0], gr.pops[k][1], gr.pops[k][2], k == 0 ? (np == 3 ? W1 : W2) : W2);
            if (q == 27 && gr.n27)
                for (int k = 0; k < gr.n27; k++) apply(gr.pops27[k][0], gr.pops27[k][1], gr.pops27[k][2], W3);
        }
    }

    // gpuSchSpecial of the rheometer (parallelPlatesRheometer.cu:139-198):
    // velocity bounce-back from the post-stream values of the node, wall
    // density = the sum of all 19 post-stream values (the five whose source
    // is an unused plane are the node's own post-collision values of the step
    // before, left by the post-collision save in step()); FRONT rotates at the
    // outer rotation, BACK is at rest. Other directions do nothing.
    // Pinned by: tests derive every incoming population as moving-wall
    // bounce-back with the rigid rotation omega (-y, x, 0) and compare.
    void bc_special(int dir, int x, int y, int z) {
        if (q != 19 || (dir != F && dir != B)) return;
        const R xNode = x - (nx - 1) / 2.0, yNode = y - (ny - 1) / 2.0;
        const R rr = std::sqrt(xNode * xNode + yNode * yNode);
        const R c = xNode / (rr), s = yNode / (rr);
        const R w_f = (R)d.outer_rotation, w_b = 0.0;
        const R W1 = w[1], W2 = w[7];
        R rho_w = PS(x, y, z, 0);
        for (int i = 1; i < 19; i++) rho_w = rho_w + PS(x, y, z, i);
        if (dir == F) {
            const R ux_w = -w_f * rr * s, uy_w = w_f * rr * c, uz_w = 0;
            PS(x, y, z, 6) = PS(x, y, z, 5) - 6 * rho_w * W1 * (uz_w);
            PS(x, y, z, 10) = PS(x, y, z, 9) - 6 * rho_w * W2 * (uz_w + ux_w);
            PS(x, y, z, 12) = PS(x, y, z, 11) - 6 * rho_w * W2 * (uz_w + uy_w);
            PS(x, y, z, 15) = PS(x, y, z, 16) - 6 * rho_w * W2 * (uz_w - ux_w);
            PS(x, y, z, 17) = PS(x, y, z, 18) - 6 * rho_w * W2 * (uz_w - uy_w);
        } else {
            const R ux_w = -w_b * rr * s, uy_w = w_b * rr * c, uz_w = 0;
            PS(x, y, z, 5) = PS(x, y, z, 6) - 6 * rho_w * W1 * (-uz_w);
            PS(x, y, z, 9) = PS(x, y, z, 10) - 6 * rho_w * W2 * (-uz_w - ux_w);
            PS(x, y, z, 11) = PS(x, y, z, 12) - 6 * rho_w * W2 * (-uz_w - uy_w);
            PS(x, y, z, 16) = PS(x, y, z, 15) - 6 * rho_w * W2 * (-uz_w + ux_w);
            PS(x, y, z, 18) = PS(x, y, z, 17) - 6 * rho_w * W2 * (-uz_w + uy_w);
        }
    }

    void apply_bc(size_t idx) {
        const int x = (int)(idx % nx), y = (int)((idx / nx) % ny), z = (int)(idx / ((size_t)nx * ny));
        const uint32_t m = map[idx];
        switch (w_scheme(m)) {
        case BOUNCE_BACK: bc_bounce_back(w_dir(m), w_geom(m), x, y, z); break;
        case VEL_BB: bc_vel_bb(w_dir(m), x, y, z, (R)d.ux_bc[w_uxi(m)], (R)d.uy_bc[w_uyi(m)], (R)d.uz_bc[w_uzi(m)]); break;
        case VEL_ZOUHE: bc_vel_zouhe(w_dir(m), x, y, z, (R)d.ux_bc[w_uxi(m)], (R)d.uy_bc[w_uyi(m)], (R)d.uz_bc[w_uzi(m)]); break;
        case PRES_ZOUHE: bc_pres_zouhe(w_dir(m), x, y, z, (R)d.rho_bc[w_rhoi(m)]); break;
        case FREE_SLIP: bc_free_slip(w_dir(m), x, y, z); break;
        case INTERP_BB: bc_interp_bb(w_interp_bits(m), w_dir(m) == S, x, y, z); break;
        case SPECIAL: bc_special(w_dir(m), x, y, z); break;
        default: break;
        }
    }

    void step(bool save, int nthreads) {
        const R FXv = (R)d.fx, FYv = (R)d.fy, FZv = (R)d.fz;
        static const int NEG_Z[] = {6, 10, 12, 15, 17, 20, 21, 24, 26};
        static const int POS_Z[] = {5, 9, 11, 16, 18, 19, 22, 23, 25};
        #pragma omp parallel for collapse(2) schedule(static) num_threads(nthreads)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++) {
                R f[27], mo[6];
                const int yp1 = (y + 1) % ny, ym1 = (ny + y - 1) % ny;
                const int zp1 = (z + 1) % (nz + 1), zm1 = ((nz + 1) + z - 1) % (nz + 1);
                for (int x = 0; x < nx; x++) {
                    const size_t k = is(x, y, z);
                    const uint32_t m = map[k];
                    if (!w_used(m)) continue;
                    for (int i = 0; i < q; i++) f[i] = pop[ip(x, y, z, i)];
                    if (!force_x.empty()) {
                        // IBM builds read the per-node force and reset it to the
                        // constant force (src/CUDA/lbm.cu:45-58)
                        collide(f, force_x[k], force_y[k], force_z[k], &omega[k], mo);
                        force_x[k] = FXv; force_y[k] = FYv; force_z[k] = FZv;
                    } else {
                        collide(f, FXv, FYv, FZv, &omega[k], mo);
                    }
                    if (save) {
                        rho[k] = mo[0]; ux[k] = mo[1]; uy[k] = mo[2]; uz[k] = mo[3];
                        if (d.rheo == 1) omega[k] = mo[4];
                    }
                    // written every step, as the kernel does (lbm.cu:237-239)
                    if (d.les_const != 0 && d.les_visc_turb) visc_turb[k] = mo[5];
                    if (w_save_post_col(m))
                        for (int i = 0; i < q; i++) pop[ip(x, y, z, i)] = f[i];
                    const int xp1 = (x + 1) % nx, xm1 = (nx + x - 1) % nx;
                    for (int i = 0; i < q; i++) {
                        const int xd = CX[i] > 0 ? xp1 : (CX[i] < 0 ? xm1 : x);
                        const int yd = CY[i] > 0 ? yp1 : (CY[i] < 0 ? ym1 : y);
                        const int zd = CZ[i] > 0 ? zp1 : (CZ[i] < 0 ? zm1 : z);
                        popAux[ip(xd, yd, zd, i)] = f[i];
                    }
                }
            }
        // ghost-plane fold (single lattice: base == next)
        const int nfold = (q == 27) ? 9 : 5;
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++)
                for (int k = 0; k < nfold; k++) {
                    popAux[ip(x, y, nz - 1, NEG_Z[k])] = popAux[ip(x, y, nz, NEG_Z[k])];
                    popAux[ip(x, y, 0, POS_Z[k])] = popAux[ip(x, y, nz, POS_Z[k])];
                }
        for (size_t k = 0; k < bc_list.size(); k++) apply_bc(bc_list[k]);
        pop.swap(popAux);
    }
};


// ---------------------------------------------------------------------------
// Immersed boundary stage for spheres (src/CUDA/IBM/ibm.cu:6-1197,
// collision/ibmCollision.cu), FP64, serial loops in the reference's order.
// ---------------------------------------------------------------------------
} // namespace

extern "C" {
struct or_ibm_desc {
    int max_iterations;                 // IBM_MAX_ITERATION
    int periodic_x, periodic_y, periodic_z;   // IBM_BC_{X,Y,Z}_PERIODIC (else wall)
    double thickness;                   // IBM_THICKNESS
    double alpha;                       // IBM_MOVEMENT_DISCRETIZATION
    double fluid_density;               // FLUID_DENSITY
    double gx, gy, gz;                  // GX, GY, GZ
    double pp_friction, pw_friction, pp_rest, pw_rest;
    double young, poisson, wall_young, wall_poisson;
    int ref_float_helpers;              // 1: float sqrt/rsqrt helpers as the reference
};
}

namespace {

constexpr int kMaxColl = 13;

struct PCenter {
    double pos[3], pos_old[3], vel[3], vel_old[3], w[3], w_avg[3], w_old[3], w_pos[3];
    double f[3], f_old[3], M[3], M_old[3];
    double I[6];   // xx yy zz xy xz yz
    double dP[3], dL[3];
    double S, radius, volume, density;
    bool movable;
    int partner[kMaxColl];
    double tang[kMaxColl][3];
    int last[kMaxColl];
    double axis[3][3];   // collision.semiAxis, semiAxis2, semiAxis3 (points)
};

// Peskin 4-point kernel (src/CUDA/IBM/ibmGlobalFunctions.h:28-48)
inline double stencil4(double x) {
    const double a = std::fabs(x);
    if (a <= 1) return (1.0 / 8.0) * (3.0 - 2.0 * a + std::sqrt(1 + 4 * a - 4 * a * a));
    if (a > 1.0 && a <= 2.0) return (1.0 / 8.0) * (5.0 - 2.0 * a - std::sqrt(-7.0 + 12.0 * a - 4.0 * a * a));
    return 0.0;
}

struct Ibm {
    Lattice<double>* L;
    or_ibm_desc d;
    std::vector<PCenter> pc;
    std::vector<double> px, py, pz, nS, nfx, nfy, nfz, ndx, ndy, ndz;
    std::vector<int> nc;
    std::vector<double> velAux[3], fAux[3];

    double pp_kn, pw_kn, pp_kt, pw_kt;

    Ibm(Lattice<double>* lat, const or_ibm_desc& desc) : L(lat), d(desc) {
        const size_t n = L->nnodes;
        for (int a = 0; a < 3; a++) { velAux[a].assign(n, 0.0); fAux[a].assign(n, 0.0); }
        L->force_x.assign(n, L->d.fx); L->force_y.assign(n, L->d.fy); L->force_z.assign(n, L->d.fz);
        // src/CUDA/IBM/ibmBoundaryCondition.h:77-96
        const double pG = d.young / (2.0 + 2.0 * d.poisson), wG = d.wall_young / (2.0 + 2.0 * d.wall_poisson);
        const double S2 = 1.41421356237309504880168872420969807856967187537;
        pp_kn = (4.0 / 3.0) / ((1 - d.poisson * d.poisson) / d.young + (1 - d.poisson * d.poisson) / d.young);
        pw_kn = (4.0 / 3.0) / ((1 - d.poisson * d.poisson) / d.young + (1 - d.wall_poisson * d.wall_poisson) / d.wall_young);
        pp_kt = 4.0 * S2 / ((2 - d.poisson) / pG + (2 - d.poisson) / pG);
        pw_kt = 4.0 * S2 / ((2 - d.poisson) / pG + (2 - d.wall_poisson) / wG);
    }

    double vlen(const double v[3]) const {
        if (d.ref_float_helpers) return (double)sqrtf((float)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]));
        return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    void vnorm(const double v[3], double o[3]) const {
        double inv;
        if (d.ref_float_helpers) inv = (double)(1.0f / sqrtf((float)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])));
        else inv = 1.0 / std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (std::isnan(inv) || std::isinf(inv)) { o[0] = o[1] = o[2] = 0; return; }
        o[0] = v[0] * inv; o[1] = v[1] * inv; o[2] = v[2] * inv;
    }

    int axis_n(int a) const { return a == 0 ? L->nx : (a == 1 ? L->ny : L->nz); }
    bool periodic(int a) const { return a == 0 ? d.periodic_x : (a == 1 ? d.periodic_y : d.periodic_z); }

    // gpuUpdateParticleOldValues (ibm.cu:995-1042); M_old is not updated there
    void old_values() {
        for (auto& c : pc) {
            for (int a = 0; a < 3; a++) {
                c.dP[a] = 0; c.dL[a] = 0;
                c.pos_old[a] = c.pos[a]; c.vel_old[a] = c.vel[a]; c.w_old[a] = c.w[a];
                c.f_old[a] = c.f[a]; c.f[a] = 0; c.M[a] = 0;
            }
        }
    }

    // gpuUpdateMacrIBM (ibm.cu:544-648): rho, u of the post-step populations
    // with the constant force; auxiliaries zeroed.
    void update_macr() {
        const int q = L->q;
        const double F[3] = {L->d.fx, L->d.fy, L->d.fz};
        for (int z = 0; z < L->nz; z++)
            for (int y = 0; y < L->ny; y++)
                for (int x = 0; x < L->nx; x++) {
                    const size_t k = L->is(x, y, z);
                    for (int a = 0; a < 3; a++) { velAux[a][k] = 0; fAux[a][k] = 0; }
                    double f[27] = {};
                    for (int i = 0; i < q; i++) f[i] = L->pop[L->ip(x, y, z, i)];
                    double r = f[0];
                    for (int i = 1; i < q; i++) r += f[i];
                    const double inv = 1 / r;
                    double j[3] = {0, 0, 0};
                    double jp[3] = {0, 0, 0}, jm[3] = {0, 0, 0};
                    for (int i = 1; i < q; i++) {
                        const int c[3] = {CX[i], CY[i], CZ[i]};
                        for (int a = 0; a < 3; a++) { if (c[a] > 0) jp[a] += f[i]; else if (c[a] < 0) jm[a] += f[i]; }
                    }
                    for (int a = 0; a < 3; a++) j[a] = (jp[a] - jm[a] + 0.5 * F[a]) * inv;
                    L->rho[k] = r; L->ux[k] = j[0]; L->uy[k] = j[1]; L->uz[k] = j[2];
                }
    }

    // ---- contacts (collision/ibmCollision.cu) ----
    static int wall_index(const double n[3]) {
        return 7 + (1 - (int)n[0]) - 2 * (1 - (int)n[1]) - 3 * (1 - (int)n[2]);
    }
    void end_collision(PCenter& c, int idx, int step) {
        if (idx >= 0 && idx <= 6) c.last[idx] = -1;
        else if (idx >= 7 && idx < kMaxColl) {
            if (step - c.last[idx] > 1) {
                c.tang[idx][0] = c.tang[idx][1] = c.tang[idx][2] = 0;
                c.last[idx] = -1; c.partner[idx] = -1;
            }
        }
    }
    int start_collision(PCenter& c, int partner, bool wall, const double n[3], int step) {
        int index = -1;
        if (wall) {
            index = wall_index(n);
            c.tang[index][0] = c.tang[index][1] = c.tang[index][2] = 0;
            c.last[index] = step;
        } else {
            for (int i = 7; i < kMaxColl; i++)
                if (c.last[i] == -1) {
                    c.partner[i] = partner; c.tang[i][0] = c.tang[i][1] = c.tang[i][2] = 0;
                    c.last[i] = step; index = i; break;
                }
        }
        return index;
    }
    void update_tang(PCenter& c, int idx, const double dsp[3], int step, double out[3]) {
        for (int a = 0; a < 3; a++) { c.tang[idx][a] += dsp[a]; out[a] = c.tang[idx][a]; }
        c.last[idx] = step;
    }
    int index_by_partner(const PCenter& c, int partner, int step) const {
        for (int i = 7; i < kMaxColl; i++)
            if (c.partner[i] == partner && step - c.last[i] <= 1) return i;
        return -1;
    }

    // sphereWallCollision (ibmCollision.cu:394-508)
    void sphere_wall(PCenter& c, const double wall_normal[3], double displacement, int step) {
        const double m_i = c.volume * c.density, r_i = c.radius;
        double n[3] = {-wall_normal[0], -wall_normal[1], -wall_normal[2]};
        const double* G = c.vel;
        const double* w_i = c.w;
        const double kn = pw_kn * std::sqrt(std::fabs(r_i));
        const double kt = pw_kt * std::sqrt(r_i) * std::sqrt(std::fabs(displacement));
        const double lr = std::log(d.pw_rest);
        const double dc = (-2.0 * lr / (std::sqrt(M_PI * M_PI + lr * lr)));
        const double dn = dc * std::sqrt(m_i * kn), dt = dc * std::sqrt(m_i * kt);
        const double f_kn = -kn * std::sqrt(std::fabs(displacement * displacement * displacement));
        const double Gn = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
        const double p25 = std::pow(std::fabs(displacement), 0.25);
        double fn[3];
        for (int a = 0; a < 3; a++) fn[a] = f_kn * n[a] - dn * Gn * n[a] * p25;
        const double f_n = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        double Gct[3];
        Gct[0] = G[0] + r_i * (w_i[1] * n[2] - w_i[2] * n[1]) - Gn * n[0];
        Gct[1] = G[1] + r_i * (w_i[2] * n[0] - w_i[0] * n[2]) - Gn * n[1];
        Gct[2] = G[2] + r_i * (w_i[0] * n[1] - w_i[1] * n[0]) - Gn * n[2];
        double mag = std::sqrt(Gct[0] * Gct[0] + Gct[1] * Gct[1] + Gct[2] * Gct[2]);
        double t[3] = {0, 0, 0};
        if (mag != 0) for (int a = 0; a < 3; a++) t[a] = Gct[a] / mag;
        const int ti = wall_index(n);
        double td[3];
        if (step - c.last[ti] > 1) {
            end_collision(c, ti, step);
            start_collision(c, ti, true, n, step);
            for (int a = 0; a < 3; a++) td[a] = Gct[a];
        } else {
            update_tang(c, ti, Gct, step, td);
        }
        double ft[3];
        for (int a = 0; a < 3; a++) ft[a] = -kt * td[a] - dt * Gct[a] * std::pow(std::fabs(td[a]), 0.25);
        mag = std::sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
        const double fnab = d.ref_float_helpers ? (double)fabsf((float)f_n) : std::fabs(f_n);
        if (mag > d.pw_friction * fnab) {
            const double mg[3] = {-Gct[0], -Gct[1], -Gct[2]};
            update_tang(c, ti, mg, step, td);
            for (int a = 0; a < 3; a++) ft[a] = -d.pw_friction * f_n * t[a];
        }
        const double fd[3] = {fn[0] + ft[0], fn[1] + ft[1], fn[2] + ft[2]};
        const double md[3] = {r_i * (n[1] * ft[2] - n[2] * ft[1]), r_i * (n[2] * ft[0] - n[0] * ft[2]),
                              r_i * (n[0] * ft[1] - n[1] * ft[0])};
        for (int a = 0; a < 3; a++) { c.f[a] += fd[a]; c.M[a] += md[a]; }
    }

    // checkCollisionWallsSphere (ibmCollision.cu:188-251): walls at 0 and N-1
    void walls(PCenter& c, int step) {
        for (int a = 0; a < 3; a++) {
            if (periodic(a)) continue;
            double nrm[3] = {0, 0, 0};
            nrm[a] = 1;
            double dist = c.pos[a] * 1.0 - 0.0;
            if (dist < c.radius) sphere_wall(c, nrm, c.radius - dist, step);
            nrm[a] = -1;
            dist = (double)(axis_n(a) - 1) + (-c.pos[a]);
            if (dist < c.radius) sphere_wall(c, nrm, c.radius - dist, step);
        }
    }

    double min_image(int a, double di, double pi, double pj) const {
        if (!periodic(a)) return di;
        const double Lx = axis_n(a);
        if (std::fabs(pi - pj) > Lx / 2.0) return pi < pj ? (pi + Lx - pj) : (pi - Lx - pj);
        return di;
    }

    // sphereSphereGap (ibmCollision.cu:1061-1088): wrap with N/2, float sqrt in the reference
    double gap(const PCenter& a, const PCenter& b) const {
        double dl[3];
        for (int k = 0; k < 3; k++) {
            dl[k] = a.pos[k] - b.pos[k];
            if (periodic(k)) {
                const double Lk = axis_n(k);
                if (dl[k] > Lk / 2.0) dl[k] -= Lk;
                if (dl[k] < -Lk / 2.0) dl[k] += Lk;
            }
        }
        const double s2 = dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2];
        const double dist = d.ref_float_helpers ? (double)sqrtf((float)s2) : std::sqrt(s2);
        return dist - (a.radius + b.radius);
    }

    // sphereSphereCollision (ibmCollision.cu:1719-1913)
    void sphere_sphere(unsigned row, PCenter& ci, PCenter& cj, int step) {
        const double r_i = ci.radius, r_j = cj.radius;
        const double m_i = ci.volume * ci.density, m_j = cj.volume * cj.density;
        double dp[3];
        for (int a = 0; a < 3; a++) dp[a] = min_image(a, ci.pos[a] - cj.pos[a], ci.pos[a], cj.pos[a]);
        const double md = std::sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
        if (md > r_i + r_j) return;
        const double n[3] = {dp[0] / md, dp[1] / md, dp[2] / md};
        const double displacement = r_i + r_j - md;
        const double G[3] = {ci.vel[0] - cj.vel[0], ci.vel[1] - cj.vel[1], ci.vel[2] - cj.vel[2]};
        const double er = 1.0 / ((r_i + r_j) / (r_i * r_j));
        const double em = 1.0 / ((m_i + m_j) / (m_i * m_j));
        const double kn = pp_kn * std::sqrt(er);
        const double kt = pp_kt * std::sqrt(er) * std::sqrt(std::fabs(displacement));
        const double lr = std::log(d.pp_rest);
        const double dc = (-2.0 * lr / (std::sqrt(M_PI * M_PI + lr * lr)));
        const double dn = dc * std::sqrt(em * kn), dt = dc * std::sqrt(em * kt);
        const double f_kn = -kn * std::sqrt(std::fabs(displacement * displacement * displacement));
        const double Gn = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
        const double p25 = std::pow(std::fabs(displacement), 0.25);
        double fn[3];
        for (int a = 0; a < 3; a++) fn[a] = f_kn * n[a] - dn * Gn * n[a] * p25;
        const double f_n = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
        const double* wi = ci.w;
        const double* wj = cj.w;
        double Gct[3];
        Gct[0] = G[0] + r_i * (wi[1] * n[2] - wi[2] * n[1]) + r_j * (wj[1] * n[2] - wj[2] * n[1]) - Gn * n[0];
        Gct[1] = G[1] + r_i * (wi[2] * n[0] - wi[0] * n[2]) + r_j * (wj[2] * n[0] - wj[0] * n[2]) - Gn * n[1];
        Gct[2] = G[2] + r_i * (wi[0] * n[1] - wi[1] * n[0]) + r_j * (wj[0] * n[1] - wj[1] * n[0]) - Gn * n[2];
        double mag = std::sqrt(Gct[0] * Gct[0] + Gct[1] * Gct[1] + Gct[2] * Gct[2]);
        double t[3] = {0, 0, 0};
        if (mag != 0) for (int a = 0; a < 3; a++) t[a] = Gct[a] / mag;
        double td[3];
        int ti = index_by_partner(ci, (int)row, step);
        if (ti == -1) {
            ti = start_collision(ci, (int)row, false, n, step);
            for (int a = 0; a < 3; a++) td[a] = Gct[a];
        } else if (step - ci.last[ti] > 1) {
            end_collision(ci, ti, step);
            ti = start_collision(ci, (int)row, false, n, step);
            for (int a = 0; a < 3; a++) td[a] = Gct[a];
        } else {
            update_tang(ci, ti, G, step, td);   // the reference accumulates G here, not G_ct
        }
        double ft[3];
        for (int a = 0; a < 3; a++) ft[a] = -kt * td[a] - dt * Gct[a] * std::pow(std::fabs(td[a]), 0.25);
        mag = std::sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
        if (mag > d.pp_friction * std::fabs(f_n)) {
            const double mg[3] = {-Gct[0], -Gct[1], -Gct[2]};
            if (ti >= 0) update_tang(ci, ti, mg, step, td);
            for (int a = 0; a < 3; a++) ft[a] = -d.pp_friction * f_n * t[a];
        }
        const double fd[3] = {fn[0] + ft[0], fn[1] + ft[1], fn[2] + ft[2]};
        const double mi[3] = {r_i * (n[1] * ft[2] - n[2] * ft[1]), r_i * (n[2] * ft[0] - n[0] * ft[2]),
                              r_i * (n[0] * ft[1] - n[1] * ft[0])};
        const double mj[3] = {r_j * (n[1] * ft[2] - n[2] * ft[1]), r_j * (n[2] * ft[0] - n[0] * ft[2]),
                              r_j * (n[0] * ft[1] - n[1] * ft[0])};
        for (int a = 0; a < 3; a++) {
            ci.f[a] += -fd[a]; ci.M[a] += mi[a];
            cj.f[a] += fd[a]; cj.M[a] += mj[a];
        }
    }

    // gpuParticlesCollisionHandler (ibmCollision.cu:2381-2448), Floyd-triangle order;
    // the reference's guard admits idx == total, which reads past the array: skipped.
    void collisions(int step) {
        const unsigned np = (unsigned)pc.size();
        const unsigned total = np * (np + 1) / 2;
        for (unsigned idx = 0; idx < total; idx++) {
            const unsigned row = (unsigned)std::ceil((-1.0 + std::sqrt((double)1 + 8 * (idx + 1))) / 2);
            const unsigned column = idx - ((row - 1) * row) / 2;
            PCenter& ci = pc[column];
            if (row == np) {
                if (!ci.movable) continue;
                walls(ci, step);
            } else {
                PCenter& cj = pc[row];
                if (!ci.movable && !cj.movable) continue;
                if (gap(ci, cj) < 0) sphere_sphere(row, ci, cj, step);
            }
        }
    }

    // gpuUpdateParticleCenterVelocityAndRotation (ibm.cu:755-897)
    void update_velocity() {
        const double al = d.alpha;
        const double g[3] = {d.gx, d.gy, d.gz};
        for (auto& c : pc) {
            if (!c.movable) continue;
            const double inv_volume = 1 / c.volume;
            for (int a = 0; a < 3; a++)
                c.vel[a] = c.vel_old[a] + (((c.f_old[a] * (1.0 - al) + c.f[a] * al) + c.dP[a]) * inv_volume +
                                           (c.density - d.fluid_density) * g[a]) / (c.density);
            const double* I = c.I;
            const double xx = I[0], yy = I[1], zz = I[2], xy = I[3], xz = I[4], yz = I[5];
            double wAux[3] = {c.w_old[0], c.w_old[1], c.w_old[2]};
            const double inv_det_neg = 1.0 / (zz * xy * xy + yy * xz * xz + xx * yz * yz - xx * yy * zz - 2 * xy * xz * yz);
            double wA[3], LM[3];
            for (int a = 0; a < 3; a++) {
                wA[a] = (c.w_old[a] * (1.0 - al) + wAux[a] * al);
                LM[a] = c.dL[a] + (c.M[a] * al + c.M_old[a] * (1.0 - al));
            }
            // I w components
            const double Iwx = xx * wA[0] + xy * wA[1] + xz * wA[2];
            const double Iwy = xy * wA[0] + yy * wA[1] + yz * wA[2];
            const double Iwz = xz * wA[0] + yz * wA[1] + zz * wA[2];
            const double Ax = LM[0] + wA[2] * Iwy - wA[1] * Iwz;
            const double Ay = LM[1] + wA[0] * Iwz - wA[2] * Iwx;
            const double Az = LM[2] + wA[1] * Iwx - wA[0] * Iwy;
            double wN[3];
            wN[0] = c.w_old[0] + ((yz * yz - yy * zz) * Ax - (xy * yz - xz * yy) * Az - (xz * yz - xy * zz) * Ay) * inv_det_neg;
            wN[1] = c.w_old[1] + ((xz * xz - xx * zz) * Ay - (xy * xz - xx * yz) * Az - (xz * yz - xy * zz) * Ax) * inv_det_neg;
            wN[2] = c.w_old[2] + ((xy * xy - xx * yy) * Az - (xy * xz - xx * yz) * Ay - (xy * yz - xz * yy) * Ax) * inv_det_neg;
            // inertia rotated by the average rotation (axis_angle_to_quart / rotate_inertia_by_quart)
            const double ang = vlen(wA);
            double ax[3];
            vnorm(wA, ax);
            const double qw = std::cos(ang / 2.0), qx = ax[0] * std::sin(ang / 2.0), qy = ax[1] * std::sin(ang / 2.0),
                         qz = ax[2] * std::sin(ang / 2.0);
            double R[3][3];
            R[0][0] = 1 - 2 * (qy * qy + qz * qz); R[0][1] = 2 * (qx * qy - qw * qz); R[0][2] = 2 * (qx * qz + qw * qy);
            R[1][0] = 2 * (qx * qy + qw * qz); R[1][1] = 1 - 2 * (qx * qx + qz * qz); R[1][2] = 2 * (qy * qz - qw * qx);
            R[2][0] = 2 * (qx * qz - qw * qy); R[2][1] = 2 * (qy * qz + qw * qx); R[2][2] = 1 - 2 * (qx * qx + qy * qy);
            const double M3[3][3] = {{xx, xy, xz}, {xy, yy, yz}, {xz, yz, zz}};
            double T[3][3], N3[3][3];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    T[i][j] = 0;
                    for (int k = 0; k < 3; k++) T[i][j] += R[i][k] * M3[k][j];
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    N3[i][j] = 0;
                    for (int k = 0; k < 3; k++) N3[i][j] += T[i][k] * R[j][k];
                }
            for (int a = 0; a < 3; a++) c.w[a] = wN[a];
            c.I[0] = N3[0][0]; c.I[1] = N3[1][1]; c.I[2] = N3[2][2];
            c.I[3] = N3[0][1]; c.I[4] = N3[0][2]; c.I[5] = N3[1][2];
        }
    }

    // gpuForceInterpolationSpread (ibm.cu:194-542), walls clip the stencil,
    // periodic axes wrap it (single GPU).
    void interp_spread() {
        const size_t nn = px.size();
        for (size_t i = 0; i < nn; i++) {
            const double pos[3] = {px[i], py[i], pz[i]};
            int base[3], mn[3], mx[3];
            for (int a = 0; a < 3; a++) {
                base[a] = (int)pos[a] - 2 + 1;
                const int Na = axis_n(a);
                if (periodic(a)) { mn[a] = 0; mx[a] = 3; }
                else {
                    mx[a] = ((base[a] + 3) < Na) ? 3 : (Na - 1 - base[a]);
                    mn[a] = (base[a] >= 0) ? 0 : -base[a];
                }
            }
            if (mx[0] <= 0 || mx[1] <= 0 || mx[2] <= 0) continue;
            if (mn[0] >= 4 || mn[1] >= 4 || mn[2] >= 4) continue;
            double sv[3][4];
            for (int a = 0; a < 3; a++)
                for (int j = mn[a]; j <= mx[a]; j++) sv[a][j] = stencil4(base[a] + j - pos[a]);
            auto node_idx = [&](int xi, int yj, int zk) -> size_t {
                int c3[3] = {base[0] + xi, base[1] + yj, base[2] + zk};
                for (int a = 0; a < 3; a++)
                    if (periodic(a)) { const int Na = axis_n(a); c3[a] = (c3[a] + Na) % Na; }
                return L->is(c3[0], c3[1], c3[2]);
            };
            double rv = 0, uxv = 0, uyv = 0, uzv = 0;
            for (int zk = mn[2]; zk <= mx[2]; zk++)
                for (int yj = mn[1]; yj <= mx[1]; yj++) {
                    const double a1 = sv[2][zk] * sv[1][yj];
                    for (int xi = mn[0]; xi <= mx[0]; xi++) {
                        const double a = a1 * sv[0][xi];
                        const size_t k = node_idx(xi, yj, zk);
                        rv += L->rho[k] * a; uxv += L->ux[k] * a; uyv += L->uy[k] * a; uzv += L->uz[k] * a;
                    }
                }
            PCenter& c = pc[nc[i]];
            double dx = px[i] - c.pos[0], dy = py[i] - c.pos[1], dz = pz[i] - c.pos[2];
            double* dd[3] = {&dx, &dy, &dz};
            for (int a = 0; a < 3; a++) {
                if (!periodic(a)) continue;
                const double Na = axis_n(a);
                if (std::fabs(*dd[a]) > Na / 2.0) *dd[a] = (*dd[a] < 0) ? (pos[a] + Na) - c.pos[a] : (pos[a] - Na) - c.pos[a];
            }
            double uc[3] = {0, 0, 0};
            if (c.movable) {
                uc[0] = c.vel[0] + (c.w[1] * dz - c.w[2] * dy);
                uc[1] = c.vel[1] + (c.w[2] * dx - c.w[0] * dz);
                uc[2] = c.vel[2] + (c.w[0] * dy - c.w[1] * dx);
            }
            const double aux = 2 * rv * nS[i] * d.thickness;
            const double dF[3] = {aux * (uxv - uc[0]), aux * (uyv - uc[1]), aux * (uzv - uc[2])};
            for (int zk = mn[2]; zk <= mx[2]; zk++)
                for (int yj = mn[1]; yj <= mx[1]; yj++) {
                    const double a1 = sv[2][zk] * sv[1][yj];
                    for (int xi = mn[0]; xi <= mx[0]; xi++) {
                        const double a = a1 * sv[0][xi];
                        const size_t k = node_idx(xi, yj, zk);
                        fAux[0][k] += -dF[0] * a; fAux[1][k] += -dF[1] * a; fAux[2][k] += -dF[2] * a;
                        const double ir = 1 / L->rho[k];
                        velAux[0][k] += 0.5 * -dF[0] * a * ir;
                        velAux[1][k] += 0.5 * -dF[1] * a * ir;
                        velAux[2][k] += 0.5 * -dF[2] * a * ir;
                    }
                }
            nfx[i] += dF[0]; nfy[i] += dF[1]; nfz[i] += dF[2];
            ndx[i] = dF[0]; ndy[i] = dF[1]; ndz[i] = dF[2];
            const double dm[3] = {dy * dF[2] - dz * dF[1], dz * dF[0] - dx * dF[2], dx * dF[1] - dy * dF[0]};
            for (int a = 0; a < 3; a++) { c.f[a] += dF[a]; c.M[a] += dm[a]; }
        }
    }

    // gpuEulerSumIBMAuxsReset (particleEulerNodesUpdate.cu:287-330)
    void euler_sum() {
        for (size_t k = 0; k < L->nnodes; k++) {
            L->ux[k] += velAux[0][k]; L->uy[k] += velAux[1][k]; L->uz[k] += velAux[2][k];
            L->force_x[k] += fAux[0][k]; L->force_y[k] += fAux[1][k]; L->force_z[k] += fAux[2][k];
            for (int a = 0; a < 3; a++) { velAux[a][k] = 0; fAux[a][k] = 0; }
        }
    }

    double wrap_pos(int a, double v) const {
        const double Na = axis_n(a);
        return std::fmod(v + Na, Na);
    }

    // gpuParticleMovement (ibm.cu:899-993). The semi-axis points are not used
    // by sphere contacts; they rotate about the centre with the node rotation
    // (including its (qi*qj - q0*qj) entry) and no guard on |w_avg| = 0, which
    // makes them NaN after a step without rotation (ibm.cu:958-991).
    void movement() {
        const double al = d.alpha;
        for (auto& c : pc) {
            if (!c.movable) continue;
            for (int a = 0; a < 3; a++) {
                const double dp = (c.vel[a] * al + c.vel_old[a] * (1.0 - al));
                if (periodic(a)) c.pos[a] = wrap_pos(a, c.pos[a] + dp);
                else c.pos[a] += dp;
            }
            for (int a = 0; a < 3; a++) {
                c.w_avg[a] = (c.w[a] * al + c.w_old[a] * (1.0 - al));
                c.w_pos[a] += c.w_avg[a];
            }
            const double wn = std::sqrt(c.w_avg[0] * c.w_avg[0] + c.w_avg[1] * c.w_avg[1] + c.w_avg[2] * c.w_avg[2]);
            const double q0 = std::cos(0.5 * wn);
            const double qi = (c.w_avg[0] / wn) * std::sin(0.5 * wn);
            const double qj = (c.w_avg[1] / wn) * std::sin(0.5 * wn);
            const double qk = (c.w_avg[2] / wn) * std::sin(0.5 * wn);
            const double t = (q0 * q0) - 0.5;
            for (auto& ax : c.axis) {
                const double x = ax[0] - c.pos_old[0], y = ax[1] - c.pos_old[1], z = ax[2] - c.pos_old[2];
                ax[0] = c.pos[0] + 2 * ((t + (qi * qi)) * x + ((qi * qj) - (q0 * qk)) * y + ((qi * qk) + (q0 * qj)) * z);
                ax[1] = c.pos[1] + 2 * (((qi * qj) + (q0 * qk)) * x + (t + (qj * qj)) * y + ((qj * qk) - (q0 * qi)) * z);
                ax[2] = c.pos[2] + 2 * (((qi * qj) - (q0 * qj)) * x + ((qj * qk) + (q0 * qi)) * y + (t + (qk * qk)) * z);
            }
        }
    }

    // gpuParticleNodeMovement (ibm.cu:1044-1197), including the reference's
    // (qi*qj - q0*qj) entry of the rotation.
    void node_movement() {
        for (size_t i = 0; i < px.size(); i++) {
            const PCenter& c = pc[nc[i]];
            if (!c.movable) continue;
            double* p3[3] = {&px[i], &py[i], &pz[i]};
            const double wn = std::sqrt(c.w_avg[0] * c.w_avg[0] + c.w_avg[1] * c.w_avg[1] + c.w_avg[2] * c.w_avg[2]);
            if (wn <= 1e-8) {
                for (int a = 0; a < 3; a++) {
                    double dd = c.pos[a] - c.pos_old[a];
                    if (periodic(a)) {
                        const double Na = axis_n(a);
                        if (std::fabs(dd) > Na / 2.0) dd = (c.pos[a] < c.pos_old[a]) ? (c.pos[a] + Na) - c.pos_old[a] : (c.pos[a] - Na) - c.pos_old[a];
                        *p3[a] = wrap_pos(a, *p3[a] + dd);
                    } else {
                        *p3[a] += dd;
                    }
                }
                continue;
            }
            double v[3];
            for (int a = 0; a < 3; a++) {
                v[a] = *p3[a] - c.pos_old[a];
                if (periodic(a)) {
                    const double Na = axis_n(a);
                    if (std::fabs(v[a]) > Na / 2.0) {
                        if (*p3[a] < c.pos_old[a]) *p3[a] += Na; else *p3[a] -= Na;
                    }
                    v[a] = *p3[a] - c.pos_old[a];
                }
            }
            const double q0 = std::cos(0.5 * wn);
            const double qi = (c.w_avg[0] / wn) * std::sin(0.5 * wn);
            const double qj = (c.w_avg[1] / wn) * std::sin(0.5 * wn);
            const double qk = (c.w_avg[2] / wn) * std::sin(0.5 * wn);
            const double t = (q0 * q0) - 0.5;
            const double nx_ = c.pos[0] + 2 * ((t + (qi * qi)) * v[0] + ((qi * qj) - (q0 * qk)) * v[1] + ((qi * qk) + (q0 * qj)) * v[2]);
            const double ny_ = c.pos[1] + 2 * (((qi * qj) + (q0 * qk)) * v[0] + (t + (qj * qj)) * v[1] + ((qj * qk) - (q0 * qi)) * v[2]);
            const double nz_ = c.pos[2] + 2 * (((qi * qj) - (q0 * qj)) * v[0] + ((qj * qk) + (q0 * qi)) * v[1] + (t + (qk * qk)) * v[2]);
            const double np3[3] = {nx_, ny_, nz_};
            for (int a = 0; a < 3; a++) *p3[a] = periodic(a) ? wrap_pos(a, np3[a]) : np3[a];
        }
    }

    void reset_node_forces() {
        std::fill(nfx.begin(), nfx.end(), 0.0); std::fill(nfy.begin(), nfy.end(), 0.0); std::fill(nfz.begin(), nfz.end(), 0.0);
        std::fill(ndx.begin(), ndx.end(), 0.0); std::fill(ndy.begin(), ndy.end(), 0.0); std::fill(ndz.begin(), ndz.end(), 0.0);
    }

    // immersedBoundaryMethod (ibm.cu:6-192)
    void stage(int step) {
        old_values();
        update_macr();
        reset_node_forces();
        collisions(step);
        update_velocity();
        for (int it = 0; it < d.max_iterations; it++) {
            interp_spread();
            update_velocity();
            euler_sum();
        }
        movement();
        node_movement();
    }
};

struct Handle {
    int prec;
    void* lat;
    Ibm* ibm = nullptr;
};

} // namespace

extern "C" {

// Node-type map words for a descriptor, x fastest ([NZ][NY][NX]).
int or_build_map(const or_desc* d, uint32_t* out) {
    size_t k = 0;
    for (int z = 0; z < d->nz; z++)
        for (int y = 0; y < d->ny; y++)
            for (int x = 0; x < d->nx; x++) out[k++] = build_word(*d, x, y, z);
    return 0;
}

void* or_create(const or_desc* d, int prec) {
    if (!d || (d->q != 19 && d->q != 27) || d->nx < 1 || d->ny < 1 || d->nz < 1) return nullptr;
    Handle* h = new Handle;
    h->prec = prec;
    if (prec == 4) h->lat = new Lattice<float>(*d);
    else h->lat = new Lattice<double>(*d);
    return h;
}

void or_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return;
    delete h->ibm;
    if (h->prec == 4) delete (Lattice<float>*)h->lat;
    else delete (Lattice<double>*)h->lat;
    delete h;
}

#define OR_DISPATCH(h, ...) \
    do { if (((Handle*)(h))->prec == 4) { auto* L = (Lattice<float>*)((Handle*)(h))->lat; __VA_ARGS__; } \
         else { auto* L = (Lattice<double>*)((Handle*)(h))->lat; __VA_ARGS__; } } while (0)

size_t or_bc_count(void* h) { size_t n = 0; OR_DISPATCH(h, n = L->bc_list.size()); return n; }
void or_bc_list(void* h, uint64_t* out) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->bc_list.size(); k++) out[k] = L->bc_list[k]);
}

void or_init(void* h, const double* rho, const double* ux, const double* uy, const double* uz) {
    OR_DISPATCH(h, L->init(rho, ux, uy, uz));
}

// Populations in the reference layout (Q planes of (NZ+1)*NY*NX), as double.
void or_get_pops(void* h, double* out) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->pop.size(); k++) out[k] = (double)L->pop[k]);
}
void or_set_pops(void* h, const double* in) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->pop.size(); k++) { L->pop[k] = in[k]; L->popAux[k] = in[k]; });
}
// Checkpoint load of both buffers (src/CUDA/simCheckpoint.cu:257-260 reads
// pop.bin into pop and popAux.bin into popAux).
void or_set_pops2(void* h, const double* pop, const double* aux) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->pop.size(); k++) { L->pop[k] = pop[k]; L->popAux[k] = aux[k]; });
}

void or_step(void* h, int nsteps, int save_last, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    OR_DISPATCH(h, for (int s = 0; s < nsteps; s++) L->step(save_last && s == nsteps - 1, nthreads));
}

// Wall seconds of nsteps (no saves), std::chrono::steady_clock around the loop.
double or_time_steps(void* h, int nsteps, int nthreads) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    const auto t0 = std::chrono::steady_clock::now();
    OR_DISPATCH(h, for (int s = 0; s < nsteps; s++) L->step(false, nthreads));
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

int or_max_threads(void) { return omp_get_max_threads(); }

// Macroscopics written by the last saving step, [NZ][NY][NX].
void or_get_macr(void* h, double* rho, double* ux, double* uy, double* uz, double* omega) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->nnodes; k++) {
        if (rho) rho[k] = L->rho[k];
        if (ux) ux[k] = L->ux[k];
        if (uy) uy[k] = L->uy[k];
        if (uz) uz[k] = L->uz[k];
        if (omega) omega[k] = L->omega[k];
    });
}
// macr.visc_turb (LES_EXPORT_VISC_TURBULENT), [NZ][NY][NX].
void or_get_visc_turb(void* h, double* out) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->nnodes; k++) out[k] = L->visc_turb[k]);
}
void or_set_omega(void* h, const double* omega) {
    OR_DISPATCH(h, for (size_t k = 0; k < L->nnodes; k++) L->omega[k] = omega[k]);
}

// Single-node collision on a 1x1x1 lattice's parameters: fin -> fout, m = (rho, ux, uy, uz, omega).
void or_collide_node(void* h, const double* fin, double fx, double fy, double fz,
                     double omega_old, double* fout, double* m) {
    OR_DISPATCH(h, {
        using R = typename std::remove_reference<decltype(L->pop[0])>::type;
        R f[27], mo[6]; R om = (R)omega_old;
        for (int i = 0; i < L->q; i++) f[i] = (R)fin[i];
        L->collide(f, (R)fx, (R)fy, (R)fz, &om, mo);
        for (int i = 0; i < L->q; i++) fout[i] = f[i];
        for (int k = 0; k < 5; k++) m[k] = mo[k];
    });
}

// Direct access to the BC schemes on a lattice whose pop (post-collision) and
// popAux (post-stream) arrays are set from the given buffers; the updated
// popAux is returned in post_stream. Used by the boundary-scheme pin tests.
// Test hook: replaces the node map (words [NZ][NY][NX]) and rebuilds the BC
// list in the same z->y->x order, so boundary schemes can be placed on faces
// no builder uses.
void or_set_map(void* h, const uint32_t* words) {
    OR_DISPATCH(h, {
        for (size_t k = 0; k < L->map.size(); k++) L->map[k] = words[k];
        L->bc_list.clear();
        for (size_t k = 0; k < L->map.size(); k++)
            if (w_used(L->map[k]) && w_scheme(L->map[k]) != BC_NULL) L->bc_list.push_back(k);
    });
}

// Test hook: the interpolated bounce-back wall distance q of a link from
// (x1, y1) towards (x2, y2) to the circle of radius r (kind 0: along x,
// 1: along y, 2: diagonal of slope +1, 3: slope -1), as bc_interp_bb uses it.
double or_ibb_q(int kind, double x1, double y1, double x2, double y2, double r) {
    using L = Lattice<double>;
    if (kind == 0) return L::q_a0(x1, y1, x2, y2, r);
    if (kind == 1) return L::q_ainf(x1, y1, x2, y2, r);
    if (kind == 2) return L::q_ap1(x1, y1, x2, y2, r);
    return L::q_am1(x1, y1, x2, y2, r);
}

void or_apply_bc_raw(void* h, double* post_stream, const double* post_col) {
    OR_DISPATCH(h, {
        for (size_t k = 0; k < L->pop.size(); k++) { L->popAux[k] = post_stream[k]; L->pop[k] = post_col[k]; }
        for (size_t k = 0; k < L->bc_list.size(); k++) L->apply_bc(L->bc_list[k]);
        for (size_t k = 0; k < L->pop.size(); k++) post_stream[k] = (double)L->popAux[k];
    });
}

} // extern "C"

// ---------------------------------------------------------------------------
// Immersed boundary entry points (FP64 lattices only)
// ---------------------------------------------------------------------------
namespace {

// makeSpherePolar (src/CUDA/IBM/structs/particle.cpp:698-995): layered polar
// mesh, equal-area rings, optional Coulomb relaxation, shifted to the centre.
int sphere_polar(double diameter, const double* center, double mesh_scale, double breugem,
                 int coulomb, double* xyz, double* S_out, int cap) {
    double r = diameter / 2.0;
    const double S_total = 4.0 * M_PI * r * r;
    r -= breugem;
    const unsigned nLayer = (unsigned)(2.0 * std::sqrt(2) * r / mesh_scale + 1.0);
    std::vector<unsigned> nNodes(nLayer + 1);
    std::vector<double> theta(nLayer + 1), zeta(nLayer + 1), S(nLayer + 1);
    unsigned total = 0;
    for (unsigned i = 0; i <= nLayer; i++) {
        theta[i] = M_PI * ((double)i / (double)nLayer - 0.5);
        nNodes[i] = (unsigned)(1.5 + std::cos(theta[i]) * nLayer * std::sqrt(3));
        total += nNodes[i];
        zeta[i] = r * std::sin(theta[i]);
    }
    if (xyz == nullptr) return (int)total;
    if ((int)total > cap) return -1;
    for (unsigned i = 0; i < nLayer; i++) S[i] = (zeta[i] + zeta[i + 1]) / 2.0 - zeta[0];
    S[nLayer] = 2 * r;
    for (unsigned i = 0; i <= nLayer; i++) S[i] = 2 * M_PI * r * S[i];
    for (unsigned i = nLayer; i > 0; i--) S[i] = S[i] - S[i - 1];
    S[0] = S[nLayer];
    std::vector<double> P(3 * (size_t)total, 0.0), A(total, 0.0);
    P[0] = 0; P[1] = 0; P[2] = r * std::sin(theta[0]); A[0] = S[0];
    unsigned node = 1;
    double phase = 0.0;
    for (unsigned i = 1; i < nLayer; i++) {
        if (i % 2 == 1) phase = phase + M_PI / nNodes[i];
        for (unsigned j = 0; j < nNodes[i]; j++) {
            P[3 * node + 0] = r * std::cos(theta[i]) * std::cos((double)j * 2.0 * M_PI / nNodes[i] + phase);
            P[3 * node + 1] = r * std::cos(theta[i]) * std::sin((double)j * 2.0 * M_PI / nNodes[i] + phase);
            P[3 * node + 2] = r * std::sin(theta[i]);
            A[node] = S[i] / nNodes[i];
            node++;
        }
    }
    // north pole (the pole layers hold one node each)
    P[3 * (total - 1) + 0] = 0; P[3 * (total - 1) + 1] = 0;
    P[3 * (total - 1) + 2] = r * std::sin(theta[nLayer]);
    A[total - 1] = S[nLayer];
    if (coulomb != 0) {
        std::vector<double> cf(3 * (size_t)total);
        const double scaleF = 0.001;
        for (int c = 0; c < coulomb; c++) {
            std::fill(cf.begin(), cf.end(), 0.0);
            for (unsigned i = 0; i < total; i++)
                for (unsigned j = i + 1; j < total; j++) {
                    const double dx = P[3 * j] - P[3 * i], dy = P[3 * j + 1] - P[3 * i + 1], dz = P[3 * j + 2] - P[3 * i + 2];
                    const double mag = dx * dx + dy * dy + dz * dz;
                    cf[3 * i] -= dx / mag; cf[3 * i + 1] -= dy / mag; cf[3 * i + 2] -= dz / mag;
                    cf[3 * j] -= -dx / mag; cf[3 * j + 1] -= -dy / mag; cf[3 * j + 2] -= -dz / mag;
                }
            for (unsigned i = 0; i < total; i++) {
                for (int a = 0; a < 3; a++) P[3 * i + a] += cf[3 * i + a] / scaleF;
                const double mag = std::sqrt(P[3 * i] * P[3 * i] + P[3 * i + 1] * P[3 * i + 1] + P[3 * i + 2] * P[3 * i + 2]);
                for (int a = 0; a < 3; a++) P[3 * i + a] *= r / mag;
            }
        }
        for (unsigned i = 0; i < total; i++) A[i] = S_total / total;
    }
    for (unsigned i = 0; i < total; i++) {
        for (int a = 0; a < 3; a++) xyz[3 * i + a] = P[3 * i + a] + center[a];
        S_out[i] = A[i];
    }
    return (int)total;
}

} // namespace

extern "C" {

int or_sphere_polar(double diameter, const double* center, double mesh_scale, double breugem, int coulomb,
                    double* xyz, double* S, int cap) {
    return sphere_polar(diameter, center, mesh_scale, breugem, coulomb, xyz, S, cap);
}

int or_ibm_enable(void* hv, const or_ibm_desc* d) {
    Handle* h = (Handle*)hv;
    if (h->prec != 8 || h->ibm) return -1;
    h->ibm = new Ibm((Lattice<double>*)h->lat, *d);
    return 0;
}

// state: pos[3], vel[3], w[3]; nodes: absolute positions xyz[3n], areas S[n]
int or_ibm_add_particle(void* hv, const double* state, double radius, double density, int movable,
                        int n_nodes, const double* xyz, const double* S) {
    Handle* h = (Handle*)hv;
    if (!h->ibm) return -1;
    Ibm& I = *h->ibm;
    PCenter c{};
    for (int a = 0; a < 3; a++) {
        c.pos[a] = c.pos_old[a] = state[a];
        c.vel[a] = c.vel_old[a] = state[3 + a];
        c.w[a] = c.w_avg[a] = c.w_old[a] = state[6 + a];
    }
    c.radius = radius;
    c.volume = radius * radius * radius * 4 * M_PI / 3;
    c.S = 4.0 * M_PI * radius * radius;
    c.density = density;
    const double Ival = 2.0 * c.volume * density * radius * radius / 5.0;
    c.I[0] = c.I[1] = c.I[2] = Ival;
    c.movable = movable != 0;
    for (int i = 0; i < kMaxColl; i++) { c.partner[i] = -1; c.last[i] = -1; }
    // semiAxis = (r, r, r) (particle.cpp:777-778); semiAxis2/3 unset there, 0 here
    for (int a = 0; a < 3; a++) { c.axis[0][a] = radius; c.axis[1][a] = 0; c.axis[2][a] = 0; }
    const int id = (int)I.pc.size();
    I.pc.push_back(c);
    for (int k = 0; k < n_nodes; k++) {
        I.px.push_back(xyz[3 * k]); I.py.push_back(xyz[3 * k + 1]); I.pz.push_back(xyz[3 * k + 2]);
        I.nS.push_back(S[k]); I.nc.push_back(id);
        I.nfx.push_back(0); I.nfy.push_back(0); I.nfz.push_back(0);
        I.ndx.push_back(0); I.ndy.push_back(0); I.ndz.push_back(0);
    }
    return id;
}

// n coupled steps: LBM step then the IBM stage labelled first_step + s
void or_ibm_run(void* hv, int nsteps, int first_step, int nthreads) {
    Handle* h = (Handle*)hv;
    Lattice<double>* L = (Lattice<double>*)h->lat;
    for (int s = 0; s < nsteps; s++) {
        L->step(false, nthreads);
        h->ibm->stage(first_step + s);
    }
}

void or_ibm_stage(void* hv, int step) { ((Handle*)hv)->ibm->stage(step); }
void or_ibm_collisions(void* hv, int step) { ((Handle*)hv)->ibm->collisions(step); }
void or_ibm_update_velocity(void* hv) { ((Handle*)hv)->ibm->update_velocity(); }

// sphere and node movement of the current velocities (the end of a stage)
void or_ibm_move(void* hv) {
    ((Handle*)hv)->ibm->movement();
    ((Handle*)hv)->ibm->node_movement();
}

int or_ibm_count(void* hv, int* n_nodes) {
    Handle* h = (Handle*)hv;
    if (n_nodes) *n_nodes = (int)h->ibm->px.size();
    return (int)h->ibm->pc.size();
}

// 48 doubles per particle: pos vel w f M I[6] pos_old vel_old w_old w_avg
// w_pos f_old M_old dP dL (3 each unless noted)
void or_ibm_get_particles(void* hv, double* out) {
    Ibm& I = *((Handle*)hv)->ibm;
    for (size_t p = 0; p < I.pc.size(); p++) {
        const PCenter& c = I.pc[p];
        double* o = out + 48 * p;
        const double* src[] = {c.pos, c.vel, c.w, c.f, c.M};
        int k = 0;
        for (auto s : src) for (int a = 0; a < 3; a++) o[k++] = s[a];
        for (int a = 0; a < 6; a++) o[k++] = c.I[a];
        const double* src2[] = {c.pos_old, c.vel_old, c.w_old, c.w_avg, c.w_pos, c.f_old, c.M_old, c.dP, c.dL};
        for (auto s : src2) for (int a = 0; a < 3; a++) o[k++] = s[a];
    }
}

void or_ibm_set_particles(void* hv, const double* in) {
    Ibm& I = *((Handle*)hv)->ibm;
    for (size_t p = 0; p < I.pc.size(); p++) {
        PCenter& c = I.pc[p];
        const double* o = in + 48 * p;
        double* dst[] = {c.pos, c.vel, c.w, c.f, c.M};
        int k = 0;
        for (auto s : dst) for (int a = 0; a < 3; a++) s[a] = o[k++];
        for (int a = 0; a < 6; a++) c.I[a] = o[k++];
        double* dst2[] = {c.pos_old, c.vel_old, c.w_old, c.w_avg, c.w_pos, c.f_old, c.M_old, c.dP, c.dL};
        for (auto s : dst2) for (int a = 0; a < 3; a++) s[a] = o[k++];
    }
}

// semi-axis points, 9 doubles per particle (semiAxis, semiAxis2, semiAxis3)
void or_ibm_get_axes(void* hv, double* out) {
    Ibm& I = *((Handle*)hv)->ibm;
    for (size_t p = 0; p < I.pc.size(); p++)
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 3; a++) out[9 * p + 3 * k + a] = I.pc[p].axis[k][a];
}

void or_ibm_get_collisions(void* hv, int* partner, double* tang, int* last) {
    Ibm& I = *((Handle*)hv)->ibm;
    for (size_t p = 0; p < I.pc.size(); p++)
        for (int i = 0; i < kMaxColl; i++) {
            partner[p * kMaxColl + i] = I.pc[p].partner[i];
            last[p * kMaxColl + i] = I.pc[p].last[i];
            for (int a = 0; a < 3; a++) tang[(p * kMaxColl + i) * 3 + a] = I.pc[p].tang[i][a];
        }
}

// node positions, accumulated forces and last force increments, [3][n] each
void or_ibm_get_nodes(void* hv, double* pos, double* f, double* df) {
    Ibm& I = *((Handle*)hv)->ibm;
    const size_t n = I.px.size();
    for (size_t k = 0; k < n; k++) {
        pos[k] = I.px[k]; pos[n + k] = I.py[k]; pos[2 * n + k] = I.pz[k];
        f[k] = I.nfx[k]; f[n + k] = I.nfy[k]; f[2 * n + k] = I.nfz[k];
        df[k] = I.ndx[k]; df[n + k] = I.ndy[k]; df[2 * n + k] = I.ndz[k];
    }
}

// per-node lattice force field (what the next collision reads), [3][nodes]
void or_ibm_get_force(void* hv, double* f) {
    Handle* h = (Handle*)hv;
    Lattice<double>* L = (Lattice<double>*)h->lat;
    const size_t n = L->nnodes;
    for (size_t k = 0; k < n; k++) { f[k] = L->force_x[k]; f[n + k] = L->force_y[k]; f[2 * n + k] = L->force_z[k]; }
}

void or_ibm_set_force(void* hv, const double* f) {
    Handle* h = (Handle*)hv;
    Lattice<double>* L = (Lattice<double>*)h->lat;
    const size_t n = L->nnodes;
    for (size_t k = 0; k < n; k++) { L->force_x[k] = f[k]; L->force_y[k] = f[n + k]; L->force_z[k] = f[2 * n + k]; }
}

// one interpolation/spread pass on the current macroscopics; returns the
// auxiliary fields [6][nodes] (velAux xyz, fAux xyz) without summing them
void or_ibm_interp_spread(void* hv, double* aux_out) {
    Ibm& I = *((Handle*)hv)->ibm;
    I.interp_spread();
    const size_t n = I.L->nnodes;
    for (int a = 0; a < 3; a++)
        for (size_t k = 0; k < n; k++) { aux_out[a * n + k] = I.velAux[a][k]; aux_out[(3 + a) * n + k] = I.fAux[a][k]; }
}

void or_ibm_update_macr(void* hv) { ((Handle*)hv)->ibm->update_macr(); }

void or_ibm_set_macr(void* hv, const double* rho, const double* ux, const double* uy, const double* uz) {
    Lattice<double>* L = (Lattice<double>*)((Handle*)hv)->lat;
    for (size_t k = 0; k < L->nnodes; k++) { L->rho[k] = rho[k]; L->ux[k] = ux[k]; L->uy[k] = uy[k]; L->uz[k] = uz[k]; }
}

double or_stencil4(double x) { return stencil4(x); }

// Cell of each sphere in the uniform cell list used by the product's contact
// search (no reference counterpart: the reference tests all pairs). Cells are
// at least 2 r_max (1 + 1e-4) wide, index floor(p / width) clamped to the grid.
void or_ibm_cells(int nx, int ny, int nz, double r_max, const double* pos, int n, int* dims, int* cells) {
    const int N[3] = {nx, ny, nz};
    const double cs = 2.0 * r_max * 1.0001;
    double w[3];
    for (int a = 0; a < 3; a++) {
        int c = cs > 0 ? (int)(N[a] / cs) : 1;
        dims[a] = c < 1 ? 1 : c;
        w[a] = (double)N[a] / dims[a];
    }
    for (int i = 0; i < n; i++) {
        int k[3];
        for (int a = 0; a < 3; a++) {
            int v = (int)std::floor(pos[3 * i + a] / w[a]);
            k[a] = v < 0 ? 0 : (v >= dims[a] ? dims[a] - 1 : v);
        }
        cells[i] = k[0] + dims[0] * (k[1] + dims[1] * k[2]);
    }
}

} // extern "C"
