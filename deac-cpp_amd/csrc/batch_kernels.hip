// Batched DEAC on gfx950: one persistent workgroup per reconstruction.
//
// Each instance runs the reference generation loop (deac.cpp:1138-1573) with
// the reference's random stream and FP64 evaluation order, so its results are
// those of deac_run (and of the reference CPU build) bit for bit.  Instances
// are independent: no grid synchronisation, and a workgroup whose instance
// stops early frees its CU for the next queued instance.
//
// Wavefront roles inside a workgroup (1024 threads, four per SIMD):
//   0-3   chain waves: model ISF E = K.S on the FP64 matrix cores.  Each
//         v_mfma_f64_16x16x4f64 continues 256 sequential fused chains by four
//         genome positions in ascending order, which is exactly the
//         reference's FMA chain (matrix_multiply_LxM_by_MxN, deac.cpp:110-118);
//         up to four rows past the last full 16-row tile run the same way on
//         v_mfma_f64_4x4x4_4b_f64, whose B operand is the 16x16x4 B register.
//   4-11  workers: stage the population chunk by chunk into LDS, rebuild the
//         genomes accepted in the previous generation and form this
//         generation's trial vectors (deac.cpp:1261-1311).
//   12    walk wave: the sequential part of the random stream (jDE redraws,
//         donor indices, mask-state skips) two generations ahead.
//   13-15 mask waves: the next generation's crossover draws from per-lane
//         sub-streams of each member's mask stream.
// The genome is processed in chunks of C positions (C * padded members = 2048
// doubles); the stages of one chunk run in consecutive steps separated by
// workgroup barriers: load(c), rebuild(c-1), trial(c-2), chains(c-3).  The
// workers issue a chunk's global loads at the top of its step and store the
// data into LDS at the bottom, so the loads fly while the step computes.
//
// Population storage: [P][Lpad] member-major, one buffer.  A generation's
// accepted trials are not written when they are selected; the next
// generation recomputes them from the same inputs (stored genomes, masks, F,
// donors) while its chunks pass through LDS, and writes them back then, so
// HBM sees one read of the population and one write per accepted genome.
#include <hip/hip_runtime.h>

#include "batch_kernels.hpp"
#include "device_common.hpp"

namespace deac {

namespace {

constexpr int kWave = 64;
constexpr int kWorker0 = 4;
constexpr int kWorkers = 8;            // 4-11: workers
constexpr int kTasks = 32 / kWorkers;  // 64-position tasks per worker and step
constexpr int kWalkWave = 12;
constexpr int kRngWave0 = 13;          // 13-15: mask waves
constexpr int kCD = kBatchChunkDoubles;
constexpr int kOffPop = 0;             // 4 buffers: genomes as stored, [Ppad][C]
constexpr int kOffNew = 4 * kCD;       // 2 buffers: rebuilt (accepted) genomes, [Ppad][C]
constexpr int kOffTrial = 6 * kCD;     // 2 buffers: trial genomes in MFMA B order
constexpr int kMaskWords = 256;        // mask words per chunk: 2 generations x Ppad x 2 candidates x C/32

enum { kPassInit = 0, kPassEvolve = 1, kPassApply = 2, kPassPre = 3 };

typedef double d4 __attribute__((ext_vector_type(4)));
// Global-memory views of the descriptor's pointers: loads and stores through
// them are global_* instructions, counted by vmcnt only, so LDS waits never
// wait for them.
typedef __attribute__((address_space(1))) double gdouble;
typedef __attribute__((address_space(1))) const double gcdouble;
typedef __attribute__((address_space(1))) uint32_t guint;
typedef __attribute__((address_space(1))) const uint64_t gcu64;
typedef __attribute__((address_space(1))) const char gchar;

// Random choices of one member in one generation (deac.cpp:1180-1240).
struct WalkRec {
    uint64_t mask_state[4];   // state before the member's crossover draws
    double cr_value[2];       // redrawn CR per sign (valid when flagged)
    double f_value[2];        // redrawn F per sign
    uint32_t mutant[3];
    uint32_t flags;           // bit 2s: CR redrawn, bit 2s+1: F redrawn
};

struct BatchShared {
    uint64_t rows[256 * 4];                  // T^L by rows
    WalkRec rec[4][kBatchMaxP];              // ring over generations
    uint64_t gen_start[4][4];                // state at the top of each generation
    uint64_t walk_state[4];                  // walk position between passes
    double fitness[kBatchMaxP];
    double cr_old[kBatchMaxP][2], f_old[kBatchMaxP][2];
    double cr_trial[2][kBatchMaxP][2], f_trial[2][kBatchMaxP][2];
    double y0t[kBatchMaxP], y0r[kBatchMaxP];  // trial value at position 0 (zero-frequency match)
    int accepted[3][kBatchMaxP];
    int flag;                                 // 0 evolve, 1 stop test fired, 2 end of the launch
    // Crossover words of the chunks in flight, [chunk % 4][generation: this,
    // previous][member][candidate][C/32], staged from global memory.
    uint32_t mring[4][kMaskWords];
};

// LDS index of trial value (member m, chunk position sl) in the order the
// MFMA B operand reads it: 4-position blocks of 16 members, lane
// (sl & 3) * 16 + (m & 15), with the member XOR-swizzled by sl & 15 so that
// both the writers (one member, 64 consecutive positions) and the readers
// (one block) touch 16 distinct bank pairs per 16 lanes.
template <int NM>
__device__ __forceinline__ int trial_index(int m, int sl)
{
    return ((sl >> 2) * NM + (m >> 4)) * 64 + (sl & 3) * 16 + ((m & 15) ^ (sl & 15));
}

// Workgroup barrier of the chunk pipeline: LDS traffic retired, global loads
// in flight are kept (each wave waits for what it must before arriving).
__device__ __forceinline__ void step_barrier()
{
    DEAC_WAIT_LDS();
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Two adjacent doubles (16-byte aligned) from global memory, one dwordx4 load.
__device__ __forceinline__ double2 load2(const gcdouble *p)
{
    double2 v;
    v.x = p[0];
    v.y = p[1];
    return v;
}

// u < cr for u = (x >> 11) * 2^-53  <=>  (x >> 11) < ceil(cr * 2^53).
__device__ __forceinline__ uint64_t cr_threshold(double cr)
{
    if (!(cr > 0.0)) return 0;
    if (cr >= 1.0) return UINT64_C(1) << 53;
    return static_cast<uint64_t>(ceil(cr * 0x1.0p53));
}

__device__ __forceinline__ double de_value(uint32_t bit, double F, double par, double x1, double x2, double x3,
                                           int allow_neg)
{
    if (!bit) return par;
    const double v = fma(F, x2 - x3, x1);  // deac.cpp:1275-1290
    return allow_neg ? v : fabs(v);
}

// s <- M s with M given as a nibble table tab[n][v] = M (v << 4n).
__device__ __forceinline__ void nibble_jump(const gcu64 *tab, uint64_t (&s)[4])
{
    uint64_t o[4] = {0, 0, 0, 0};
#pragma unroll
    for (int w = 0; w < 4; w++) {
#pragma unroll 4
        for (int n = 0; n < 16; n++) {
            const int v = static_cast<int>((s[w] >> (4 * n)) & 15);
            const gcu64 *e = tab + ((w * 16 + n) * 16 + v) * 4;
            o[0] ^= e[0];
            o[1] ^= e[1];
            o[2] ^= e[2];
            o[3] ^= e[3];
        }
    }
    for (int w = 0; w < 4; w++) s[w] = o[w];
}

// One item of the sequential walk of generation `gen`: items 0..P-1 are the
// jDE redraws of member i (CR+, F+, CR-, F-: deac.cpp:1180-1213), items
// P..2P-1 the donor indices of member i (deac.cpp:200-216) followed by a skip
// of its nsign*G crossover draws.  Every lane carries the same state.
__device__ void walk_item(const BatchDesc &d, BatchShared &sh, uint64_t (&s)[4], long gen, int item, int lane)
{
    const int slot = static_cast<int>(gen & 3);
    const int P = d.P;
    if (item == 0 && lane == 0)
        for (int w = 0; w < 4; w++) sh.gen_start[slot][w] = s[w];
    if (item < P) {
        uint32_t flags = 0;
        double crv[2] = {0.0, 0.0}, fv[2] = {0.0, 0.0};
        for (int sg = 0; sg < d.nsign; sg++) {
            if (to_unit(xnext(s)) < d.tau_cr) {
                flags |= 1u << (2 * sg);
                crv[sg] = to_unit(xnext(s));
            }
            if (to_unit(xnext(s)) < d.tau_f) {
                flags |= 2u << (2 * sg);
                fv[sg] = 2.0 * to_unit(xnext(s));
            }
        }
        if (lane == 0) {
            WalkRec &r = sh.rec[slot][item];
            r.flags = flags;
            r.cr_value[0] = crv[0];
            r.cr_value[1] = crv[1];
            r.f_value[0] = fv[0];
            r.f_value[1] = fv[1];
        }
    } else {
        const uint32_t i0 = static_cast<uint32_t>(item - P), Pu = static_cast<uint32_t>(P);
        uint32_t m0 = i0, m1 = i0, m2 = i0;
        while (m0 == i0) m0 = mod_p(xnext(s), d.mod_magic, Pu);
        while (m1 == i0 || m1 == m0) m1 = mod_p(xnext(s), d.mod_magic, Pu);
        while (m2 == i0 || m2 == m0 || m2 == m1) m2 = mod_p(xnext(s), d.mod_magic, Pu);
        if (lane == 0) {
            WalkRec &r = sh.rec[slot][i0];
            r.mutant[0] = m0;
            r.mutant[1] = m1;
            r.mutant[2] = m2;
            for (int w = 0; w < 4; w++) r.mask_state[w] = s[w];
        }
        // s <- T^L s: lane l forms output bits l, 64+l, 128+l, 192+l as
        // parities of (row & s); ballots assemble the words.
        uint64_t out[4];
        for (int w = 0; w < 4; w++) {
            const uint64_t *r = sh.rows + (w * 64 + lane) * 4;
            const uint64_t x = (r[0] & s[0]) ^ (r[1] & s[1]) ^ (r[2] & s[2]) ^ (r[3] & s[3]);
            out[w] = __ballot(parity64(x));
        }
        for (int w = 0; w < 4; w++) s[w] = out[w];
    }
}

// Crossover-mask generator of one worker lane: a sub-stream of Lsub draws of
// one member's mask stream, reached by a nibble-table jump.  Each draw sets
// two candidate bits: the member's CR if its trial is rejected in the
// generation running now, and the trial's CR if it is accepted (a redrawn CR
// makes both the same).  The next generation picks one by the selection.
struct MaskGen {
    uint64_t s[4];
    uint64_t ta, tb;      // thresholds of the two candidates at the current position
    uint64_t na, nb;      // thresholds after the sign change (negative frequencies)
    uint32_t wa, wb;
    int pos, end, flip;   // flip: position of the sign change (or past the end)
    guint *out;
};

template <int NM>
__device__ void mask_setup(MaskGen &mg, const BatchDesc &d, const BatchShared &sh, long mgen, bool evolving, long g,
                           int wl)
{
    constexpr int Ppad = 16 * NM;
    const int m = wl / d.S, k = wl % d.S;
    mg.pos = mg.end = 0;
    mg.wa = mg.wb = 0;
    mg.flip = 0;
    mg.ta = mg.tb = mg.na = mg.nb = 0;
    mg.out = (guint *)d.masks;
    if (m >= d.P) return;
    const WalkRec &r = sh.rec[mgen & 3][m];
    uint64_t thr[2][2];
    for (int sg = 0; sg < 2; sg++) {
        double a = 0.0, b = 0.0;
        if (sg < d.nsign) {
            const bool redraw = (r.flags >> (2 * sg)) & 1;
            a = redraw ? r.cr_value[sg] : sh.cr_old[m][sg];
            b = redraw ? r.cr_value[sg] : (evolving ? sh.cr_trial[g & 1][m][sg] : a);
        }
        thr[0][sg] = cr_threshold(a);
        thr[1][sg] = cr_threshold(b);
    }
    uint64_t s[4] = {r.mask_state[0], r.mask_state[1], r.mask_state[2], r.mask_state[3]};
    if (k > 0) nibble_jump((const gcu64 *)d.sub_tab + static_cast<size_t>(k - 1) * (64 * 16 * 4), s);
    for (int w = 0; w < 4; w++) mg.s[w] = s[w];
    mg.pos = min(k * d.Lsub, d.L);
    mg.end = min(mg.pos + d.Lsub, d.L);
    mg.flip = d.nsign == 2 ? d.G : d.L;
    const bool neg = mg.pos >= mg.flip;
    mg.ta = neg ? thr[0][1] : thr[0][0];
    mg.tb = neg ? thr[1][1] : thr[1][0];
    mg.na = thr[0][1];
    mg.nb = thr[1][1];
    mg.out = (guint *)d.masks + (static_cast<size_t>(mgen % 3) * Ppad + m) * 2 * d.Lw;
}

// Up to q draws (q a multiple of 8).  Bits enter each candidate word
// most-significant first (w = 2w + bit); a finished word is bit-reversed so
// that bit j belongs to draw j of the word.  Sub-streams start on word
// boundaries, so 8-draw groups never straddle a word except at the stream's
// end; the sign change of the thresholds is handled per draw only in the
// group that contains it.
__device__ __forceinline__ void mask_draws(MaskGen &mg, int q, int Lw)
{
    const int n = min(q, mg.end - mg.pos);
    for (int u = 0; u < n; u += 8) {
        const int cnt = min(8, n - u);
        if (cnt == 8 && (mg.pos >= mg.flip || mg.pos + 8 <= mg.flip)) {
            if (mg.pos >= mg.flip) {
                mg.ta = mg.na;
                mg.tb = mg.nb;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t r = xnext(mg.s) >> 11;
                mg.wa = mg.wa * 2u + (r < mg.ta ? 1u : 0u);
                mg.wb = mg.wb * 2u + (r < mg.tb ? 1u : 0u);
            }
            mg.pos += 8;
        } else {
            for (int k = 0; k < cnt; k++) {
                if (mg.pos == mg.flip) {
                    mg.ta = mg.na;
                    mg.tb = mg.nb;
                }
                const uint64_t r = xnext(mg.s) >> 11;
                mg.wa = mg.wa * 2u + (r < mg.ta ? 1u : 0u);
                mg.wb = mg.wb * 2u + (r < mg.tb ? 1u : 0u);
                mg.pos++;
            }
        }
        if ((mg.pos & 31) == 0 || mg.pos == mg.end) {
            const int nb = ((mg.pos - 1) & 31) + 1;  // draws in this word
            const int w = (mg.pos - 1) >> 5;
            mg.out[w] = __brev(mg.wa) >> (32 - nb);
            mg.out[Lw + w] = __brev(mg.wb) >> (32 - nb);
            mg.wa = mg.wb = 0;
        }
    }
}

// One pass over the genome in chunks.  kind: kPassInit (model ISF of the
// stored population), kPassEvolve (generation g), kPassApply (write back the
// genomes accepted in generation g), kPassPre (masks only).  mgen: the
// generation whose masks the workers draw (-1: none); wgen: the generation
// the walk wave walks (-1: none).
template <int NT, int NM, int REM>
__device__ __forceinline__ void run_pass(const BatchDesc &d, BatchShared &sh, double *lds, int kind, long g, long mgen,
                                         long wgen)
{
    constexpr int C = 128 / NM, Ppad = 16 * NM;
    // kernel-tile prefetch depth in pairs of 4-position blocks (divides C/8)
    constexpr int KD = NT == 1 ? (REM ? 4 : 8) : (NT == 2 ? 4 : 2);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    const int nchunk = d.nchunk;
    const bool chains = kind == kPassInit || kind == kPassEvolve;
    const int nsteps = kind == kPassApply ? nchunk + 1 : nchunk + 3;

    if (wave < kWorker0) {
        // ------------------------------------------------ chain waves
        const gcdouble *kt = (const gcdouble *)d.ktile;
        const int nkb2 = d.Lpad / 8;
        // Tiles past ntt are computed from a clamped (valid) tile and
        // discarded: straight-line loads and MFMAs keep the prefetch waits
        // exact.
        int tile[NT], tload[NT];
        bool tv[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) {
            tile[i] = wave + 4 * i;
            tv[i] = chains && tile[i] < d.ntt;
            tload[i] = min(tile[i], max(d.ntt - 1, 0));
        }
        int boff[4];
        {
            const int q = lane >> 4, j = lane & 15;
#pragma unroll
            for (int u = 0; u < 4; u++) boff[u] = q * 16 + (j ^ ((4 * u + q) & 15));
        }
        d4 acc[NT][NM];
#pragma unroll
        for (int i = 0; i < NT; i++)
#pragma unroll
            for (int mt = 0; mt < NM; mt++) acc[i][mt] = d4{0.0, 0.0, 0.0, 0.0};
        double2 kr[KD][NT];
        // Loads address a wave-uniform block base (scalar registers) plus the
        // lane's 16-byte offset.
        const gcdouble *ktw[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) ktw[i] = kt + static_cast<size_t>(tload[i]) * nkb2 * (2 * kWave);
        const int lo = 2 * lane;
#pragma unroll
        for (int p = 0; p < KD; p++)
#pragma unroll
            for (int i = 0; i < NT; i++) kr[p][i] = load2(ktw[i] + static_cast<size_t>(min(p, nkb2 - 1)) * (2 * kWave) + lo);
        // Rows past the last full tile (REM): one 4x4x4 MFMA per 4 positions
        // for member tile rmt = wave % NM; waves >= NM repeat a tile's work
        // (discarded) to keep the code straight-line.
        const gcdouble *krm = (const gcdouble *)d.krem;
        const int rmt = wave % NM;
        const uint64_t rsel = rmt ? ~UINT64_C(0) : UINT64_C(0);
        double accr = 0.0;
        double2 krr[KD];
        if (REM) {
#pragma unroll
            for (int p = 0; p < KD; p++) krr[p] = load2(krm + static_cast<size_t>(min(p, nkb2 - 1)) * (2 * kWave) + lo);
        }
        // Steps 0-2 fill the pipeline, steps 3 .. nchunk+2 run the chains of
        // chunk c-3; the hot loop has no branch, so the waits for the K
        // prefetch are exact across its iterations.
        const int nwork = chains ? nchunk : 0;
        for (int c = 0; c < min(3, nsteps); c++) step_barrier();
        for (int cm = 0; cm < nwork; cm++) {
            {
                const double *tb = lds + kOffTrial + (cm & 1) * kCD;
                // B operands one block pair ahead of the MFMAs that use them
                double bc[2][NM], bn[2][NM];
#pragma unroll
                for (int e = 0; e < 2; e++)
#pragma unroll
                    for (int mt = 0; mt < NM; mt++) bc[e][mt] = tb[(e * NM + mt) * kWave + boff[e]];
#pragma unroll
                for (int jj = 0; jj < C / 8; jj++) {
                    const int sl = jj % KD;
                    const int g2 = cm * (C / 8) + jj;
                    if (jj + 1 < C / 8) {
#pragma unroll
                        for (int e = 0; e < 2; e++) {
                            const int kbl = 2 * (jj + 1) + e;
#pragma unroll
                            for (int mt = 0; mt < NM; mt++) bn[e][mt] = tb[(kbl * NM + mt) * kWave + boff[kbl & 3]];
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 2; e++) {
#pragma unroll
                        for (int i = 0; i < NT; i++) {
                            const double a = e ? kr[sl][i].y : kr[sl][i].x;
#pragma unroll
                            for (int mt = 0; mt < NM; mt++)
                                acc[i][mt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bc[e][mt], acc[i][mt], 0, 0, 0);
                        }
                        if (REM) {
                            // B register of member tile rmt by a bit select:
                            // no branch (exact prefetch waits) and no indexed
                            // register array (which would go through scratch)
                            const double ar = e ? krr[sl].y : krr[sl].x;
                            const uint64_t b0 = __builtin_bit_cast(uint64_t, bc[e][0]);
                            const uint64_t b1 = __builtin_bit_cast(uint64_t, bc[e][NM - 1]);
                            const double br = __builtin_bit_cast(double, (b1 & rsel) | (b0 & ~rsel));
                            accr = __builtin_amdgcn_mfma_f64_4x4x4f64(ar, br, accr, 0, 0, 0);
                        }
                    }
#pragma unroll
                    for (int e = 0; e < 2; e++)
#pragma unroll
                        for (int mt = 0; mt < NM; mt++) bc[e][mt] = bn[e][mt];
                    const int gn = min(g2 + KD, nkb2 - 1);
#pragma unroll
                    for (int i = 0; i < NT; i++) kr[sl][i] = load2(ktw[i] + static_cast<size_t>(gn) * (2 * kWave) + lo);
                    if (REM) krr[sl] = load2(krm + static_cast<size_t>(gn) * (2 * kWave) + lo);
                    __builtin_amdgcn_sched_barrier(0);  // issue the refill here, a full step ahead of its use
                }
            }
            step_barrier();
        }
        for (int c = min(3, nsteps) + nwork; c < nsteps; c++) step_barrier();
        __syncthreads();
        // E[t][m] into LDS (the chunk buffers are free now): lane l holds
        // rows 4v + l/16 of the tile, column l%16.
        if (chains) {
#pragma unroll
            for (int i = 0; i < NT; i++)
                if (tv[i])
#pragma unroll
                    for (int mt = 0; mt < NM; mt++) {
                        const int m = mt * 16 + (lane & 15);
                        if (m < d.P)
#pragma unroll
                            for (int v = 0; v < 4; v++) {
                                const int t = tile[i] * 16 + 4 * v + (lane >> 4);
                                lds[t * Ppad + m] = acc[i][mt][v];
                            }
                    }
            // 4x4x4 output: lane l holds row l>>4 of block (l>>2)&3, column
            // l&3, i.e. member rmt*16 + (l & 15).
            if (REM && wave < NM) {
                const int m = rmt * 16 + (lane & 15), r = lane >> 4;
                if (m < d.P && r < d.nrem) lds[(d.ntt * 16 + r) * Ppad + m] = accr;
            }
        }
    } else if (wave < kWalkWave) {
        // ------------------------------------------------ workers
        const int wk = wave - kWorker0;
        // Descriptor fields held in scalar registers for the whole pass
        // (readfirstlane keeps them from being re-loaded from memory, which
        // would make every task wait on the scalar cache).
        const int nP = __builtin_amdgcn_readfirstlane(d.P), nG = __builtin_amdgcn_readfirstlane(d.G);
        const int nL = __builtin_amdgcn_readfirstlane(d.L), nLpad = __builtin_amdgcn_readfirstlane(d.Lpad);
        const int n2 = __builtin_amdgcn_readfirstlane(d.nsign) == 2 ? 1 : 0;
        const int aneg = __builtin_amdgcn_readfirstlane(d.allow_neg);
        const int nLw = __builtin_amdgcn_readfirstlane(d.Lw);
        constexpr int kSegs = C / 64;
        constexpr int kWpc = C / 32;  // mask words per member and chunk
        const bool rebuild = kind == kPassEvolve || kind == kPassApply;
        const long pg = kind == kPassEvolve ? g - 1 : g;      // generation whose accepted trials are rebuilt
        const int pslot = static_cast<int>((pg + 3) % 3);
        const int pprev = static_cast<int>((pg + 2) % 3);     // generation pg-1
        const int gslot = static_cast<int>((g + 3) % 3);
        const int aslot = static_cast<int>((g + 2) % 3);      // generation g-1
        gdouble *gpop = (gdouble *)d.pop;
        // Per-pass task descriptors: donors r1 | r2<<5 | r3<<10, mask
        // candidate <<15, operand from the rebuilt buffer (parent, r1, r2, r3)
        // <<16..19, active <<20.  Lane j holds trial task j, lane kTasks+j
        // rebuild task j; readlane makes them wave-uniform where used.
        uint32_t dtab = 0;
        if (lane < 2 * kTasks) {
            const int j = lane % kTasks;
            const int t = wk + kWorkers * j, m = t / kSegs;
            // The walk record is read only when the task is active: the
            // record of the generation before a launch's first one was never
            // written in that launch.
            if (lane < kTasks) {
                if (kind == kPassEvolve && m < nP) {
                    const WalkRec &r = sh.rec[g & 3][m];
                    const uint32_t cand = sh.accepted[aslot][m] & 1;
                    uint32_t td = (r.mutant[0] & 31) | ((r.mutant[1] & 31) << 5) | ((r.mutant[2] & 31) << 10) |
                                  (cand << 15) | (cand << 16) | (1u << 20);
                    for (int k = 0; k < 3; k++)
                        td |= static_cast<uint32_t>(sh.accepted[aslot][r.mutant[k] & 31] & 1) << (17 + k);
                    dtab = td;
                }
            } else if (rebuild && pg >= 0 && m < nP && sh.accepted[pslot][m]) {
                const WalkRec &r = sh.rec[pg & 3][m];
                dtab = (r.mutant[0] & 31) | ((r.mutant[1] & 31) << 5) | ((r.mutant[2] & 31) << 10) |
                       (static_cast<uint32_t>(sh.accepted[pprev][m] & 1) << 15) | (1u << 20);
            }
        }
        // load(c): the stored genomes of chunk c ([Ppad][C], 16 rows of 128
        // doubles, rows wk and wk+8 for this wave; padded members copy the
        // last member's row) and, for workers 0-3, its crossover words (this
        // generation's and the rebuilt one's).  Issued at the top of step c
        // into registers, written to LDS at its end: plain loads, so no LDS
        // read in between waits for them (an LDS DMA would make every LDS
        // read wait for the copy in flight).
        // Addresses: a wave-uniform chunk base plus per-lane 32-bit byte
        // offsets (the population of one instance stays below 2 GiB).
        constexpr int kRows = 16 / kWorkers;
        uint32_t poff[kRows], moff = 0;
#pragma unroll
        for (int u = 0; u < kRows; u++) {
            const int od = (wk + kWorkers * u) * 128 + lane * 2;
            const int m = min(od / C, nP - 1), x = od % C;
            poff[u] = static_cast<uint32_t>(m * nLpad + x) * 8u;
        }
        if (wk < kMaskWords / kWave) {
            const int e = wk * kWave + lane;
            const int w = e % kWpc, cand = (e / kWpc) & 1, m = (e / (2 * kWpc)) % Ppad;
            const int slot = e / (2 * kWpc * Ppad) == 0 ? gslot : pslot;
            moff = static_cast<uint32_t>(((slot * Ppad + m) * 2 + cand) * nLw + w) * 4u;
        }
        double2 pv[kRows];
        uint32_t mw = 0;
        for (int c = 0; c < nsteps; c++) {
            // The lane index passes through an empty asm each step, so the
            // per-task offsets derived from it are recomputed here rather
            // than held in registers across the loop (which spills them).
            int ln = lane;
            asm volatile("" : "+v"(ln));
            const bool ld = kind != kPassPre && c < nchunk;
            if (ld) {
                const gchar *pb = (const gchar *)d.pop + static_cast<size_t>(c) * C * 8;
#pragma unroll
                for (int u = 0; u < kRows; u++) pv[u] = load2((const gcdouble *)(pb + poff[u]));
                if (wk < kMaskWords / kWave)
                    mw = *(const guint *)((const gchar *)d.masks + static_cast<size_t>(c) * kWpc * 4 + moff);
            }
            // rebuild(c-1): genomes accepted in generation pg, then trial(c-2):
            // this generation's trial vectors (or, for the initial evaluation,
            // the stored genomes) in MFMA B order.  Each phase loads the
            // operands, F and mask words of all its tasks at once (one LDS
            // round trip), then computes; operands are addressed by 32-bit
            // offsets into the dynamic LDS array.
            const int c1 = c - 1;
            if (rebuild && pg >= 0 && c1 >= 0 && c1 < nchunk) {
                const int popo = kOffPop + (c1 % 4) * kCD;
                const int newo = kOffNew + (c1 & 1) * kCD;
                const uint32_t *mrb = sh.mring[c1 % 4] + Ppad * 2 * kWpc;
                double v4[kTasks][4], Fv[kTasks];
                uint32_t w4[kTasks];
#pragma unroll
                for (int j = 0; j < kTasks; j++) {
                    const uint32_t ds = __builtin_amdgcn_readlane(dtab, kTasks + j);
                    if (!((ds >> 20) & 1)) continue;  // most members were not accepted
                    const int t = wk + kWorkers * j, m = min(t / kSegs, nP - 1), seg = t % kSegs;
                    const int sl = seg * 64 + ln, s = c1 * C + sl;
                    v4[j][0] = lds[popo + m * C + sl];
                    v4[j][1] = lds[popo + static_cast<int>(ds & 31) * C + sl];
                    v4[j][2] = lds[popo + static_cast<int>((ds >> 5) & 31) * C + sl];
                    v4[j][3] = lds[popo + static_cast<int>((ds >> 10) & 31) * C + sl];
                    w4[j] = mrb[(m * 2 + static_cast<int>((ds >> 15) & 1)) * kWpc + (sl >> 5)];
                    Fv[j] = sh.f_trial[pg & 1][m][(n2 && s >= nG) ? 1 : 0];
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < kTasks; j++) {
                    const uint32_t ds = __builtin_amdgcn_readlane(dtab, kTasks + j);
                    if (!((ds >> 20) & 1)) continue;
                    const int t = wk + kWorkers * j, m = t / kSegs, seg = t % kSegs;
                    const int sl = seg * 64 + ln, s = c1 * C + sl;
                    double v = de_value((w4[j] >> (s & 31)) & 1, Fv[j], v4[j][0], v4[j][1], v4[j][2], v4[j][3],
                                        aneg);
                    if (n2 && s == nG) {  // deac.cpp:1305-1308
                        if (c1 == 0) {
                            const int wb = (m * 2 + static_cast<int>((ds >> 15) & 1)) * kWpc;
                            v = de_value(mrb[wb] & 1, sh.f_trial[pg & 1][m][0], lds[popo + m * C],
                                         lds[popo + static_cast<int>(ds & 31) * C],
                                         lds[popo + static_cast<int>((ds >> 5) & 31) * C],
                                         lds[popo + static_cast<int>((ds >> 10) & 31) * C], aneg);
                        } else {
                            v = sh.y0r[m];
                        }
                    }
                    if (s == 0) sh.y0r[m] = v;
                    if (s >= nL) v = 0.0;
                    lds[newo + m * C + sl] = v;
                    if (s < nL) (gpop + (static_cast<size_t>(m) * nLpad + c1 * C + seg * 64))[ln] = v;
                }
            }
            const int c2 = c - 2;
            if (chains && c2 >= 0 && c2 < nchunk) {
                const int popo = kOffPop + (c2 % 4) * kCD;
                const int newo = kOffNew + (c2 & 1) * kCD;
                double *tb = lds + kOffTrial + (c2 & 1) * kCD;
                const uint32_t *mrt = sh.mring[c2 % 4];
                if (kind == kPassInit) {
#pragma unroll
                    for (int j = 0; j < kTasks; j++) {
                        const int t = wk + kWorkers * j, m = t / kSegs, seg = t % kSegs;
                        if (m >= nP) continue;
                        const int sl = seg * 64 + ln, s = c2 * C + sl;
                        tb[trial_index<NM>(m, sl)] = s < nL ? lds[popo + m * C + sl] : 0.0;
                    }
                } else {
                    double v4[kTasks][4], Fv[kTasks];
                    uint32_t w4[kTasks];
#pragma unroll
                    for (int j = 0; j < kTasks; j++) {
                        const uint32_t ds = __builtin_amdgcn_readlane(dtab, j);
                        const int t = wk + kWorkers * j, m = min(t / kSegs, nP - 1), seg = t % kSegs;
                        const int sl = seg * 64 + ln, s = c2 * C + sl;
                        v4[j][0] = lds[(((ds >> 16) & 1) ? newo : popo) + m * C + sl];
#pragma unroll
                        for (int k = 0; k < 3; k++)
                            v4[j][k + 1] = lds[(((ds >> (17 + k)) & 1) ? newo : popo) +
                                               static_cast<int>((ds >> (5 * k)) & 31) * C + sl];
                        w4[j] = mrt[(m * 2 + static_cast<int>((ds >> 15) & 1)) * kWpc + (sl >> 5)];
                        Fv[j] = sh.f_trial[g & 1][m][(n2 && s >= nG) ? 1 : 0];
                    }
                    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                    for (int j = 0; j < kTasks; j++) {
                        const uint32_t ds = __builtin_amdgcn_readlane(dtab, j);
                        if (!((ds >> 20) & 1)) continue;
                        const int t = wk + kWorkers * j, m = t / kSegs, seg = t % kSegs;
                        const int sl = seg * 64 + ln, s = c2 * C + sl;
                        double v = de_value((w4[j] >> (s & 31)) & 1, Fv[j], v4[j][0], v4[j][1], v4[j][2], v4[j][3],
                                            aneg);
                        if (n2 && s == nG) {
                            if (c2 == 0) {
                                const int wb = (m * 2 + static_cast<int>((ds >> 15) & 1)) * kWpc;
                                double x[4];
                                x[0] = lds[(((ds >> 16) & 1) ? newo : popo) + m * C];
                                for (int k = 0; k < 3; k++)
                                    x[k + 1] = lds[(((ds >> (17 + k)) & 1) ? newo : popo) +
                                                   static_cast<int>((ds >> (5 * k)) & 31) * C];
                                v = de_value(mrt[wb] & 1, sh.f_trial[g & 1][m][0], x[0], x[1], x[2], x[3], aneg);
                            } else {
                                v = sh.y0t[m];
                            }
                        }
                        if (s == 0) sh.y0t[m] = v;
                        if (s >= nL) v = 0.0;
                        tb[trial_index<NM>(m, sl)] = v;
                    }
                }
            }
            // load(c) lands: the chunk's genomes and crossover words go from
            // registers into LDS for the next step's rebuild
            if (ld) {
#pragma unroll
                for (int u = 0; u < kRows; u++)
                    *reinterpret_cast<double2 *>(lds + kOffPop + (c % 4) * kCD + (wk + kWorkers * u) * 128 + 2 * ln) =
                        pv[u];
                if (wk < kMaskWords / kWave) sh.mring[c % 4][wk * kWave + ln] = mw;
            }
            step_barrier();
        }
        __syncthreads();
    } else if (wave >= kRngWave0) {
        // ------------------------------------------------ mask waves
        // The next generation's crossover draws, Lsub per lane per pass.
        MaskGen mg;
        const bool draws = mgen >= 0;
        const int q = draws ? ((d.Lsub + nsteps - 1) / nsteps + 7) / 8 * 8 : 0;
        if (draws) mask_setup<NM>(mg, d, sh, mgen, kind == kPassEvolve, g, (wave - kRngWave0) * kWave + lane);
        for (int c = 0; c < nsteps; c++) {
            if (draws) mask_draws(mg, q, d.Lw);
            step_barrier();
        }
        DEAC_WAIT_VM(0);
        __syncthreads();
    } else {
        // ------------------------------------------------ walk wave
        const int items = 2 * d.P;
        const int per = wgen >= 0 ? (items + nsteps - 1) / nsteps : 0;
        uint64_t ws[4] = {sh.walk_state[0], sh.walk_state[1], sh.walk_state[2], sh.walk_state[3]};
        int next = 0;
        for (int c = 0; c < nsteps; c++) {
            const int stop = min(items, (c + 1) * per);
            for (; next < stop; next++) walk_item(d, sh, ws, wgen, next, lane);
            step_barrier();
        }
        if (lane == 0)
            for (int w = 0; w < 4; w++) sh.walk_state[w] = ws[w];
        __syncthreads();
    }
    __syncthreads();
}

}  // namespace

template <int NT, int NM, int REM>
__global__ __launch_bounds__(kBatchThreads) void batch_kernel(const BatchDesc *__restrict__ descs)
{
    constexpr int Ppad = 16 * NM;
    extern __shared__ __attribute__((aligned(16))) double lds[];  // 16-byte aligned: ds_*_b128 at full rate
    __shared__ BatchShared sh;
    const BatchDesc &d = descs[blockIdx.x];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    const int P = d.P;
    const gcdouble *obs = (const gcdouble *)d.obs, *sig = (const gcdouble *)d.sig;

    for (int k = threadIdx.x; k < 256 * 4; k += kBatchThreads) sh.rows[k] = ((const gcu64 *)d.walk_rows)[k];
    for (int k = threadIdx.x; k < 2 * kCD; k += kBatchThreads) lds[kOffTrial + k] = 0.0;  // padded members stay 0
    if (threadIdx.x < kBatchMaxP) {
        const int m = threadIdx.x;
        const gcdouble *fit = (const gcdouble *)d.fitness, *cr = (const gcdouble *)d.cr_old,
                       *fw = (const gcdouble *)d.f_old;
        sh.fitness[m] = m < P ? fit[m] : 0.0;
        for (int sg = 0; sg < 2; sg++) {
            sh.cr_old[m][sg] = m < P ? cr[m * 2 + sg] : 0.0;
            sh.f_old[m][sg] = m < P ? fw[m * 2 + sg] : 0.0;
        }
        for (int k = 0; k < 3; k++) sh.accepted[k][m] = 0;
    }
    __syncthreads();

    // Prologue: the walk of the first generation whose masks are drawn.
    const long g0 = d.gen0;
    if (wave == kWalkWave) {
        const gcu64 *rng = (const gcu64 *)d.rng;
        uint64_t ws[4] = {rng[0], rng[1], rng[2], rng[3]};
        for (int item = 0; item < 2 * P; item++) walk_item(d, sh, ws, g0, item, lane);
        if (lane == 0)
            for (int w = 0; w < 4; w++) sh.walk_state[w] = ws[w];
    }
    __syncthreads();

    // Pass sequence: [initial evaluation | masks of gen0], then one pass per
    // generation, then the write-back of the last generation's acceptances.
    int kind = d.init ? kPassInit : kPassPre;
    long g = g0;      // generation at the top of the loop
    long last = -1;   // last generation evolved in this launch
    const int tlast = (d.T & 1) ? d.T - 1 : -1;  // fused element of the chi^2 sum
    for (;;) {
        const long pg = kind == kPassEvolve ? g : kind == kPassApply ? last : g0 - 1;
        const long mgen = kind == kPassEvolve ? g + 1 : kind == kPassApply ? -1 : g0;
        const long wgen = kind == kPassEvolve ? g + 2 : kind == kPassApply ? -1 : g0 + 1;
        run_pass<NT, NM, REM>(d, sh, lds, kind, pg, mgen, wgen);
        if (kind == kPassApply) break;

        if (kind == kPassInit || kind == kPassEvolve) {
            // Fitness (deac.cpp:1539-1555; initial fitness deac.cpp:957-976)
            // and greedy selection (deac.cpp:1556-1572).
            if (wave == kWorker0 && lane < Ppad) {
                int acc = 0;
                if (lane < P) {
                    double chi = 0.0;
                    for (int t = 0; t < d.T; t++) {
                        const double dv = (obs[t] - lds[t * Ppad + lane]) / sig[t];
                        chi = t == tlast ? fma(dv, dv, chi) : chi + dv * dv;
                    }
                    const double fit = chi / static_cast<double>(d.T);
                    if (kind == kPassInit) {
                        sh.fitness[lane] = fit;
                    } else if (fit <= sh.fitness[lane]) {
                        acc = 1;
                        sh.fitness[lane] = fit;
                        for (int sg = 0; sg < 2; sg++) {
                            sh.cr_old[lane][sg] = sh.cr_trial[g & 1][lane][sg];
                            sh.f_old[lane][sg] = sh.f_trial[g & 1][lane][sg];
                        }
                    }
                }
                if (kind == kPassEvolve) sh.accepted[g % 3][lane] = acc;
            }
            if (kind == kPassEvolve) {
                last = g;
                g++;
            }
            __syncthreads();
        }

        // Top of generation g: best fitness, statistics, stop test
        // (deac.cpp:1145-1166); this generation's CR and F (deac.cpp:1180-1213).
        if (wave == kWorker0) {
            if (lane == 0) {
                int flag = 0;
                if (g >= d.gen_end) {
                    flag = 2;
                } else {
                    double mn = sh.fitness[0];
                    for (int i = 0; i < P; i++)
                        if (sh.fitness[i] < mn) mn = sh.fitness[i];
                    if (d.stats) {
                        gdouble *st = (gdouble *)d.stats;
                        double sum = 0.0, sq = 0.0;
                        const int p4 = P & ~3;
                        for (int i = 0; i < P; i++) {
                            const double v = sh.fitness[i];
                            sum += v;
                            sq = i < p4 ? sq + v * v : fma(v, v, sq);
                        }
                        st[g] = sum / static_cast<double>(P);
                        st[d.stats_len + g] = mn;
                        st[2 * d.stats_len + g] = sq / static_cast<double>(P);
                    }
                    if (mn <= d.stop) flag = 1;
                }
                sh.flag = flag;
            }
            if (lane < P) {
                const WalkRec &r = sh.rec[g & 3][lane];
                for (int sg = 0; sg < 2; sg++) {
                    sh.cr_trial[g & 1][lane][sg] = (r.flags >> (2 * sg)) & 1 ? r.cr_value[sg] : sh.cr_old[lane][sg];
                    sh.f_trial[g & 1][lane][sg] = (r.flags >> (2 * sg + 1)) & 1 ? r.f_value[sg] : sh.f_old[lane][sg];
                }
            }
        }
        __syncthreads();
        if (sh.flag) {
            if (last < 0) break;
            kind = kPassApply;  // write back the genomes accepted in the last generation
        } else {
            kind = kPassEvolve;
        }
    }

    if (threadIdx.x < P) {
        const int m = threadIdx.x;
        gdouble *fit = (gdouble *)d.fitness, *cr = (gdouble *)d.cr_old, *fw = (gdouble *)d.f_old;
        fit[m] = sh.fitness[m];
        for (int sg = 0; sg < 2; sg++) {
            cr[m * 2 + sg] = sh.cr_old[m][sg];
            fw[m * 2 + sg] = sh.f_old[m][sg];
        }
    }
    if (threadIdx.x == 0) {
        __attribute__((address_space(1))) uint64_t *rng = (__attribute__((address_space(1))) uint64_t *)d.rng;
        __attribute__((address_space(1))) int *status = (__attribute__((address_space(1))) int *)d.status;
        for (int w = 0; w < 4; w++) rng[w] = sh.gen_start[g & 3][w];
        status[0] = sh.flag == 1 ? 1 : 0;
        status[1] = static_cast<int>(g);
    }
}

namespace {

template <int NT, int NM, int REM>
hipError_t launch_one(const BatchDesc *descs, int count, hipStream_t s)
{
    constexpr size_t lds_bytes = static_cast<size_t>(kBatchRegionDoubles) * sizeof(double);
    static bool configured = false;
    if (!configured) {
        const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&batch_kernel<NT, NM, REM>),
                                                 hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes));
        if (e != hipSuccess) return e;
        configured = true;
    }
    hipLaunchKernelGGL((batch_kernel<NT, NM, REM>), dim3(count), dim3(kBatchThreads), lds_bytes, s, descs);
    return hipGetLastError();
}

template <int NM, int REM>
hipError_t launch_nt(const BatchDesc *descs, int count, int nt, hipStream_t s)
{
    if (nt == 1) return launch_one<1, NM, REM>(descs, count, s);
    if (nt == 2) return launch_one<2, NM, REM>(descs, count, s);
    if (nt == 4) return launch_one<4, NM, REM>(descs, count, s);
    return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_batch(const BatchDesc *descs, int count, int nt, int nm, int rem, hipStream_t s)
{
    if (nm == 1) return rem ? launch_nt<1, 1>(descs, count, nt, s) : launch_nt<1, 0>(descs, count, nt, s);
    if (nm == 2) return rem ? launch_nt<2, 1>(descs, count, nt, s) : launch_nt<2, 0>(descs, count, nt, s);
    return hipErrorInvalidValue;
}

}  // namespace deac
