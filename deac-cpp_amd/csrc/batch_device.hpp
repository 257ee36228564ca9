// Batched DEAC on gfx950: one persistent workgroup per reconstruction (or
// per up to four small-population reconstructions of one problem, stacked
// along the member rows).  Device code shared by the kernel's
// instantiations: batch_kernels.hip (exact FP64), batch_mixed_kernels.hip
// (mixed-precision K.S) and batch_fold_kernels.hip (the detailed-balance
// variant, whose trial contraction is a fold of rounded products).
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
// workgroup barriers: load(c), rebuild(c-1), write-back(c-2), trial(c-3),
// chains(c-4).  The workers issue a chunk's global loads at the top of its
// step and store the data into LDS at the bottom, so the loads fly while the
// step computes.
//
// Population storage: [P][Lpad] member-major per instance (stacked
// instances at byte offsets pop_row0[q] from one base).  A generation's
// accepted trials are not written when they are selected; the next
// generation recomputes them from the same inputs (stored genomes, masks, F,
// donors) while its chunks pass through LDS and writes them over the chunk in
// place (LDS and HBM), so HBM sees one read of the population and one write
// per accepted genome, and the trial vectors read a single chunk buffer.
//
// Precision of the model ISF (template parameter PREC, a deac_precision):
//   FP64     the chain waves run the reference's FMA chains exactly (above);
//   FP32,    delta form: the parent's model ISF E_i is kept in FP64 (epar),
//   BF16X3   and a trial's is E_i + K.(trial_i - parent_i), the difference
//            taken in FP64 and contracted on the FP32 matrix cores
//            (v_mfma_f32_16x16x4f32) or as three bfloat16 products per
//            element (v_mfma_f32_16x16x32_bf16, K and the difference split
//            into hi + lo), with K scaled per timeslice by a power of two and
//            FP32 accumulation; chi^2, selection and everything else stay
//            FP64.  The difference is zero wherever the crossover keeps the
//            parent, and it shrinks as the population converges, so the
//            contraction's error is relative to |trial - parent|, not to
//            |trial|.  The mixed kernel never evaluates a population from
//            scratch: the host runs the FP64 kernel in evaluation mode
//            (mode 1: fitness and epar of the stored population, exactly)
//            before the first launch and after every launch slice.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "batch_kernels.hpp"
#include "device_common.hpp"

namespace deac {

namespace {

constexpr int kWave = 64;
constexpr int kWorker0 = 4;
constexpr int kWorkers = 8;            // 4-11: workers
constexpr int kT = 16 / kWorkers;      // chunk rows (of 128 doubles) per worker
constexpr int kWalkWave = 12;
constexpr int kRngWave0 = 13;          // 13-15: mask waves
constexpr int kCD = kBatchChunkDoubles;
constexpr int kOffPop = 0;             // 4 buffers: genomes of chunks in flight, [Ppad][C]
constexpr int kOffTrial = 4 * kCD;     // 2 buffers: trial genomes in MFMA B order
// Mixed kernel with moment constraints: the workers' partial moment sums of
// a pass, [member][C/2 lanes][2], at the end of the chunk region (past the
// model ISF [T][Ppad] that the chain waves write there after the pass)
constexpr int kPartDoubles = 2048;
constexpr int kOffPart = 6 * kCD - kPartDoubles;
constexpr int kMaskWords = 256;        // mask words per chunk: 2 generations x Ppad x 2 candidates x C/32
constexpr int kMaxT = 4 * 4 * 16 + kBatchMaxRem;  // timeslices: 16 full tiles (4 per chain wave) + remainder rows

enum { kPassInit = 0, kPassEvolve = 1, kPassApply = 2, kPassPre = 3, kPassNorm = 4, kPassNone = 5 };
enum { kPrecFp64 = 0, kPrecFp32 = 1, kPrecBf16 = 2 };  // DEAC_PRECISION_*
enum { kPrecFold = 3 };  // FP64 throughout, the trials' contraction a fold of rounded products (bdsf)
// model ISF values of a mixed-kernel workgroup: T <= 260 at 32 members, T <= 132 at 64
constexpr int kMixedIsf = kMaxT * 32 > 132 * kBatchMaxP ? kMaxT * 32 : 132 * kBatchMaxP;
constexpr int kEk = (kMixedIsf + kBatchThreads - 1) / kBatchThreads;  // chi^2 terms per thread (mixed)

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));
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
    uint64_t gen_start[4][kBatchPackMax][4]; // each stream's state at the top of each generation
    uint64_t walk_state[kBatchPackMax][4];   // each stream's walk position between passes
    double obs[kMaxT], sig[kMaxT];           // ISF and its errors
    double kscale[kMaxT];                    // 2^ex[t]: K row t is stored scaled by 2^-ex[t] (mixed K.S)
    double fitness[kBatchMaxP];
    double cr_old[kBatchMaxP][2], f_old[kBatchMaxP][2];
    double cr_trial[2][kBatchMaxP][2];
    alignas(16) double f_trial[2][kBatchMaxP][2];   // read as pairs
    double y0t[kBatchMaxP], y0r[kBatchMaxP];  // trial value at position 0 (zero-frequency match)
    double mom[kBatchMaxP][2];                // first and third moments of the pass's genomes (FP64 kernel)
    double norm[kBatchMaxP];                  // zeroth moment of each trial (renormalization)
    double nfm[kBatchMaxP];                   // -1st moment of each trial's model ISF (bdsf)
    double rscale[2][kBatchMaxP];             // renormalization factor of each trial, by generation parity
    int accepted[3][kBatchMaxP];
    int flag;                                 // 0 evolve, 1 stop test fired, 2 end of the launch
    int qstop[kBatchPackMax];                 // instances whose stop test fired
    // Crossover words of the chunks in flight, [chunk % 4][generation: this,
    // previous][member][candidate][C/32], staged from global memory.
    uint32_t mring[4][kMaskWords];
};

// LDS index of trial value (member m, chunk position sl) in the order the
// MFMA B operand reads it: 4-position blocks of 16 members, lane
// (sl & 3) * 16 + (m & 15), with the member XOR-swizzled by (sl >> 1) & 15
// so that both the writers (16 lanes: one member, positions sl, sl + 2, ..,
// sl + 30) and the readers (16 lanes: one position of a block) touch 16
// distinct bank pairs.  Position sl + 1 of an even sl lies 16 further on.
template <int NM>
__device__ __forceinline__ int trial_index(int m, int sl)
{
    return ((sl >> 2) * NM + (m >> 4)) * 64 + (sl & 3) * 16 + ((m & 15) ^ ((sl >> 1) & 15));
}

// Workgroup barrier of the chunk pipeline: LDS traffic retired, global loads
// in flight are kept (each wave waits for what it must before arriving).
#ifdef DEAC_ROLE_TIMERS
// Profiling build only (make ROLE_TIMERS=1): per wavefront of workgroup 0,
// the shader-clock cycles spent waiting in the pipeline's barriers and in
// the whole launch, and the generations evolved (deac_batch_role_cycles).
__shared__ unsigned long long g_wait[kBatchWaves];
__device__ unsigned long long g_role[kBatchWaves][3];
#endif

__device__ __forceinline__ void step_barrier()
{
    DEAC_WAIT_LDS();
    asm volatile("" ::: "memory");
#ifdef DEAC_ROLE_TIMERS
    const unsigned long long t0 = __builtin_amdgcn_s_memtime();
#endif
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
#ifdef DEAC_ROLE_TIMERS
    const unsigned long long t1 = __builtin_amdgcn_s_memtime();
    if ((threadIdx.x & 63) == 0) g_wait[threadIdx.x >> 6] += t1 - t0;
#endif
}

// A pointer held in scalar registers: its two halves made wave-uniform, so
// the compiler keeps it rather than re-reading the descriptor in loops.
template <typename T>
__device__ __forceinline__ T *uniform_ptr(T *p)
{
    const uint64_t v = reinterpret_cast<uint64_t>(p);
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
    return reinterpret_cast<T *>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Two adjacent doubles (16-byte aligned) from global memory, one dwordx4
// load, marked non-temporal: the population streams through once per
// generation and should not displace the kernel tiles every instance
// re-reads from L2.
__device__ __forceinline__ double2 load2_stream(const gcdouble *p)
{
    double2 v;
    v.x = __builtin_nontemporal_load(p);
    v.y = __builtin_nontemporal_load(p + 1);
    return v;
}

// A buffer resource over `bytes` bytes from p (raw buffer, word 3 of gfx9)
// and a load of two adjacent doubles through it at byte offset vo (per lane)
// + so (wave-uniform): one buffer_load_dwordx4 whose address needs no vector
// arithmetic.  Reads past `bytes` return zeros.
typedef __amdgpu_buffer_rsrc_t rsrc_t;
typedef double dv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ rsrc_t make_rsrc(const double *p, int bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), 0, bytes, 0x00020000);
}

__device__ __forceinline__ double2 buffer_load2(rsrc_t r, uint32_t vo, int so)
{
    const dv2 v = __builtin_bit_cast(dv2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
    double2 d;
    d.x = v.x;
    d.y = v.y;
    return d;
}

__device__ __forceinline__ u4 buffer_load4(rsrc_t r, uint32_t vo, int so)
{
    return __builtin_bit_cast(u4, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
}

// Two FP32 values as packed bfloat16, rounded to nearest even: a in the low
// half, b in the high half (v_cvt_pk_bf16_f32).
__device__ __forceinline__ uint32_t pack_bf16(float a, float b)
{
    uint32_t r;
    asm("v_cvt_pk_bf16_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}

// The mixed kernel's trial buffer (one per chunk in flight) holds the
// differences dS = trial - parent in the contraction's operand order.
//   FP32:   float at trial_index (the FP64 layout, 4-byte elements);
//   BF16X3: per 32-position block b and member m, 64 bytes of hi values then
//           (a plane further on) 64 bytes of lo values; the four 16-byte
//           groups of 8 positions are rotated by m/4, so that the 16 lanes
//           of one MFMA B read (16 members, one group) hit distinct banks.
constexpr int kBfLoPlane = kBatchChunkDoubles * 2;  // bytes from the hi plane to the lo plane

template <int NM>
__device__ __forceinline__ int bf_trial_offset(int m, int k)  // byte offset, k = position in the chunk
{
    constexpr int Ppad = 16 * NM;
    return ((k >> 5) * Ppad + m) * 64 + ((((k >> 3) & 3) + (m >> 2)) & 3) * 16 + (k & 7) * 2;
}

// Lane l's double, read by every lane (l wave-uniform): two v_readlane_b32.
__device__ __forceinline__ double readlane_f64(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(b), l);
    const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(b >> 32), l);
    return __builtin_bit_cast(double, (static_cast<uint64_t>(hi) << 32) | lo);
}

// u < cr for u = (x >> 11) * 2^-53  <=>  (x >> 11) < ceil(cr * 2^53).
__device__ __forceinline__ uint64_t cr_threshold(double cr)
{
    if (!(cr > 0.0)) return 0;
    if (cr >= 1.0) return UINT64_C(1) << 53;
    return static_cast<uint64_t>(ceil(cr * 0x1.0p53));
}

// A DE value F (x2 - x3) + x1 (deac.cpp:1275-1290) made non-negative by
// clearing the sign bit, or kept (absmask all ones) for allow_negative.
__device__ __forceinline__ double de_abs(double v, uint32_t absmask)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    return __builtin_bit_cast(double, b & ((static_cast<uint64_t>(absmask) << 32) | 0xffffffffu));
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
__device__ void walk_item(const BatchDesc &d, BatchShared &sh, uint64_t (&s)[4], long gen, int q, int item, int lane)
{
    // item j of instance q's stream: j < Pi the jDE draws of member j, then
    // member j - Pi's donors and mask-stream skip; members are numbered in
    // the workgroup (q Pi + local), donors drawn within the instance
    const int slot = static_cast<int>(gen & 3);
    const int Pi = d.Pi, base = q * Pi;
    if (item == 0 && lane == 0)
        for (int w = 0; w < 4; w++) sh.gen_start[slot][q][w] = s[w];
    if (item < Pi) {
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
            WalkRec &r = sh.rec[slot][base + item];
            r.flags = flags;
            r.cr_value[0] = crv[0];
            r.cr_value[1] = crv[1];
            r.f_value[0] = fv[0];
            r.f_value[1] = fv[1];
        }
    } else {
        const uint32_t i0 = static_cast<uint32_t>(item - Pi), Pu = static_cast<uint32_t>(Pi);
        uint32_t m0 = i0, m1 = i0, m2 = i0;
        while (m0 == i0) m0 = mod_p(xnext(s), d.mod_magic, Pu);
        while (m1 == i0 || m1 == m0) m1 = mod_p(xnext(s), d.mod_magic, Pu);
        while (m2 == i0 || m2 == m0 || m2 == m1) m2 = mod_p(xnext(s), d.mod_magic, Pu);
        if (lane == 0) {
            WalkRec &r = sh.rec[slot][base + i0];
            r.mutant[0] = base + m0;
            r.mutant[1] = base + m1;
            r.mutant[2] = base + m2;
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
//
// A draw x passes the candidate threshold T = ceil(CR 2^53) when
// (x >> 11) < T, i.e. x <= U = T 2^11 - 1 (mod 2^64) for 1 <= T <= 2^53; a
// zero threshold never passes, which the group mask k (low 8 bits clear)
// enforces after each group of 8 draws.
struct MaskGen {
    uint64_t s[4];        // xoshiro256+ state
    uint64_t ua, ub;      // U of the two candidates at the current position
    uint64_t una, unb;    // U after the sign change (negative frequencies)
    uint32_t ka, kb, kna, knb;
    uint32_t wa, wb;
    int pos, end, flip;   // flip: position of the sign change (or past the end)
    guint *out;
};

__device__ __forceinline__ uint64_t u_of_threshold(uint64_t t) { return t * 2048u - 1u; }
__device__ __forceinline__ uint32_t k_of_threshold(uint64_t t) { return t ? 0xffffffffu : 0xffffff00u; }

// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96).  Written as asm so the
// compiler keeps the three-input form instead of re-associating the xors of
// consecutive generator steps into two-input ones.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "v"(c));
    return r;
}

__device__ __forceinline__ uint64_t xor3_64(uint64_t a, uint64_t b, uint64_t c)
{
    const uint32_t lo = xor3(static_cast<uint32_t>(a), static_cast<uint32_t>(b), static_cast<uint32_t>(c));
    const uint32_t hi = xor3(static_cast<uint32_t>(a >> 32), static_cast<uint32_t>(b >> 32), static_cast<uint32_t>(c >> 32));
    return (static_cast<uint64_t>(hi) << 32) | lo;
}

// One xoshiro256+ step (rng.hpp:25-39) written over the old state:
// s2' = s2^s0^t, s1' = s1^s2^s0, s0' = s0^s3^s1, s3' = rotl(s3^s1, 45).
// Returns s0 + s3 of the old state.
__device__ __forceinline__ uint64_t xnext_mask(uint64_t (&s)[4])
{
    const uint64_t s0 = s[0], s1 = s[1], s2 = s[2], s3 = s[3];
    const uint64_t out = s0 + s3;
    const uint64_t e = s3 ^ s1;
    s[2] = xor3_64(s2, s0, s1 << 17);
    s[1] = xor3_64(s1, s2, s0);
    s[0] = s0 ^ e;
    const uint32_t lo = static_cast<uint32_t>(e), hi = static_cast<uint32_t>(e >> 32);
    s[3] = (static_cast<uint64_t>(__builtin_amdgcn_alignbit(lo, hi, 19)) << 32) | __builtin_amdgcn_alignbit(hi, lo, 19);
    return out;
}

// w = 2 w + [lane bit of mask]: one v_addc_co_u32 with the compare's lane mask
// as carry-in.
__device__ __forceinline__ uint32_t shift_in(uint32_t w, uint64_t lanemask)
{
    uint32_t r;
    uint64_t carry;
    asm("v_addc_co_u32 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry) : "v"(w), "s"(lanemask));
    return r;
}

template <int NM>
__device__ void mask_setup(MaskGen &mg, const BatchDesc &d, const BatchShared &sh, long mgen, bool evolving, long g,
                           int wl)
{
    constexpr int Ppad = 16 * NM;
    const int m = wl / d.S, k = wl % d.S;
    mg.pos = mg.end = 0;
    mg.wa = mg.wb = 0;
    mg.flip = 0;
    mg.ua = mg.ub = mg.una = mg.unb = 0;
    mg.ka = mg.kb = mg.kna = mg.knb = 0;
    for (int w = 0; w < 4; w++) mg.s[w] = 0;
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
    const int sg = mg.pos >= mg.flip ? 1 : 0;
    mg.ua = u_of_threshold(thr[0][sg]);
    mg.ub = u_of_threshold(thr[1][sg]);
    mg.ka = k_of_threshold(thr[0][sg]);
    mg.kb = k_of_threshold(thr[1][sg]);
    mg.una = u_of_threshold(thr[0][1]);
    mg.unb = u_of_threshold(thr[1][1]);
    mg.kna = k_of_threshold(thr[0][1]);
    mg.knb = k_of_threshold(thr[1][1]);
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
                mg.ua = mg.una;
                mg.ub = mg.unb;
                mg.ka = mg.kna;
                mg.kb = mg.knb;
            }
#pragma unroll
            for (int k = 0; k < 8; k++) {
                const uint64_t x = xnext_mask(mg.s);
                mg.wa = shift_in(mg.wa, __builtin_amdgcn_ballot_w64(x <= mg.ua));
                mg.wb = shift_in(mg.wb, __builtin_amdgcn_ballot_w64(x <= mg.ub));
            }
            mg.wa &= mg.ka;
            mg.wb &= mg.kb;
            mg.pos += 8;
        } else {
            for (int k = 0; k < cnt; k++) {
                if (mg.pos == mg.flip) {
                    mg.ua = mg.una;
                    mg.ub = mg.unb;
                    mg.ka = mg.kna;
                    mg.kb = mg.knb;
                }
                const uint64_t x = xnext_mask(mg.s);
                mg.wa = mg.wa * 2u + (x <= mg.ua ? mg.ka & 1u : 0u);
                mg.wb = mg.wb * 2u + (x <= mg.ub ? mg.kb & 1u : 0u);
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

// Population staging by the mixed kernel's chain waves.  Chunk c's rows wave, wave + 4,
// + 8, + 12 of the [Ppad][C] chunk buffer (two positions per lane; padded
// members read the last member) and its crossover words (this generation's
// and the previous one's, two candidates each: 256 words) are loaded at the
// bottom of step c - 1, after chunk c - 1 has landed, and stored into LDS at
// the bottom of step c.  The loads fly across a barrier and a whole step,
// beside the chain waves' own kernel-tile stream, and the workers' steps do
// not wait on HBM.
template <int NM>
struct Stager {
    static constexpr int C = 128 / NM, Ppad = 16 * NM, kWpc = C / 32;
    const char *popb, *maskb;
    uint32_t poff[4], moff;
    double2 pv[4];
    uint32_t mw;
    int nload;

    __device__ __forceinline__ void init(const BatchDesc &d, int kind, long g, int nchunk, int wave, int lane)
    {
        const int nP = d.P, nLpad = d.Lpad, nLw = d.Lw;
        popb = reinterpret_cast<const char *>(uniform_ptr(d.pop));
        maskb = reinterpret_cast<const char *>(uniform_ptr(d.masks));
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int own = (wave + 4 * r) * 128 + 2 * lane;
            poff[r] = d.pop_row0[0] + static_cast<uint32_t>(min(own / C, nP - 1) * nLpad + own % C) * 8u;
        }
        const long pg = kind == kPassEvolve ? g - 1 : g;
        const int gslot = static_cast<int>((g + 3) % 3), pslot = static_cast<int>((pg + 3) % 3);
        const int e = wave * kWave + lane;
        const int w = e % kWpc, cand = (e / kWpc) & 1, m = (e / (2 * kWpc)) % Ppad;
        const int slot = e / (2 * kWpc * Ppad) == 0 ? gslot : pslot;
        moff = static_cast<uint32_t>(((slot * Ppad + m) * 2 + cand) * nLw + w) * 4u;
        nload = kind != kPassPre && kind != kPassNone ? nchunk : 0;
        mw = 0;
    }
    __device__ __forceinline__ void issue(int c)
    {
        const gchar *pb = (const gchar *)popb + static_cast<size_t>(c) * C * 8;
#pragma unroll
        for (int r = 0; r < 4; r++) pv[r] = load2_stream((const gcdouble *)(pb + poff[r]));
        mw = *(const guint *)((const gchar *)maskb + static_cast<size_t>(c) * kWpc * 4 + moff);
    }
    __device__ __forceinline__ void start()
    {
        if (0 < nload) issue(0);
    }
    // at the bottom of step c, before its barrier
    __device__ __forceinline__ void step(int c, BatchShared &sh, double *lds, int wave, int lane)
    {
        if (c < nload) {
            double *buf = lds + kOffPop + (c % 4) * kCD + wave * 128 + 2 * lane;
#pragma unroll
            for (int r = 0; r < 4; r++) *reinterpret_cast<double2 *>(buf + r * 512) = pv[r];
            sh.mring[c % 4][wave * kWave + lane] = mw;
        }
        if (c + 1 < nload) issue(c + 1);
    }
};

// Chain waves of the mixed kernel: dE = K.dS for this wave's pairs (kernel
// tile, member tile), NQ of them.  With two member tiles, waves 0 and 1 take
// the even kernel tiles 0, 2, 4, .. and waves 2 and 3 the odd ones, each wave
// one member tile (its parity), so every wave reads one B operand per block;
// with one member tile, wave w takes tiles w, w + 4, ..; with four (33 to 64
// members, T <= 132), wave w takes member tile w and every kernel tile.  The
// kernel rows past
// T are zero rows of the last tile (no remainder chains), and pairs past the
// last tile compute on it and are discarded.
//
// K (scaled per row, see BatchShared::kscale) streams through a buffer
// resource, one load unit per lane at a time: FP32: 16 positions, four
// floats (one per 4-position block); BF16X3: one 32-position block, 16
// bytes of hi and 16 of lo.  The ring of KD units per pair is refilled one
// unit-ring ahead.  Output: dE (unscaled) into LDS as doubles, [t][Ppad].
template <int PREC, int NQ, int NM>
__device__ __forceinline__ void mixed_chains(const BatchDesc &d, BatchShared &sh, double *lds, int kind, long g,
                                             bool chains, int nchunk, int nsteps, int wave, int lane)
{
    constexpr int C = 128 / NM, Ppad = 16 * NM;
    constexpr bool BF = PREC == kPrecBf16;
    constexpr int UPC = BF ? C / 32 : C / 16;             // load units per chunk
    constexpr int RPU = BF ? 8 : 4;                       // VGPRs per unit and pair
    constexpr int KDW = 64 / (NQ * RPU);                  // ring depth that fits ~64 VGPRs
    constexpr int KD = KDW >= UPC ? UPC : KDW >= 4 ? 4 : KDW >= 2 ? 2 : 1;  // divides UPC
    static_assert(UPC % KD == 0, "the ring depth divides the units of a chunk");
    const int ntm = d.ntm;
    const int mt = NM == 4 ? wave : NM == 2 ? (wave & 1) : 0;
    int tq[NQ];
    bool tv[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const int t = NM == 4 ? q : NM == 2 ? (wave >> 1) + 2 * q : wave + 4 * q;
        tv[q] = chains && t < ntm;
        tq[q] = min(t, ntm - 1);
    }
    const int units = d.Lpad / (BF ? 32 : 16);
    const int tile_bytes = d.Lpad * 64;                   // 16 rows x Lpad positions x 4 bytes
    rsrc_t kr[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++)
        kr[q] = __builtin_amdgcn_make_buffer_rsrc(const_cast<char *>(static_cast<const char *>(d.kmix)) +
                                                      static_cast<size_t>(tq[q]) * tile_bytes,
                                                  0, tile_bytes, 0x00020000);
    const uint32_t lo_ = (BF ? 32u : 16u) * lane;
    constexpr int unit_bytes = BF ? 2048 : 1024;
    f4 acc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    int boff[8];
    {
        const int qq = lane >> 4, j = lane & 15;
#pragma unroll
        for (int u = 0; u < 8; u++) boff[u] = qq * 16 + (j ^ ((2 * u + (qq >> 1)) & 15));
    }
    const int bfoff = (mt * 16 + (lane & 15)) * 64 + ((((lane >> 4) + ((lane & 15) >> 2)) & 3) * 16);
    const int nwork = chains ? nchunk : 0;
    u4 ka[KD][NQ], kl[KD][NQ];
#pragma unroll
    for (int p = 0; p < KD; p++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const int so = min(p, units - 1) * unit_bytes;
            ka[p][q] = buffer_load4(kr[q], lo_, so);
            if (BF) kl[p][q] = buffer_load4(kr[q], lo_ + 16u, so);
        }
    Stager<NM> st;
    st.init(d, kind, g, nchunk, wave, lane);
    st.start();
    for (int c = 0; c < min(4, nsteps); c++) {
        st.step(c, sh, lds, wave, lane);
        step_barrier();
    }
    for (int cm = 0; cm < nwork; cm++) {
        const int c = cm + 4;
        const char *tb = reinterpret_cast<const char *>(lds + kOffTrial + (cm & 1) * kCD);
#pragma unroll
        for (int jj = 0; jj < UPC; jj++) {
            const int sl = jj % KD;
            if constexpr (BF) {
                const bf8 bh = *reinterpret_cast<const bf8 *>(tb + jj * Ppad * 64 + bfoff);
                const bf8 bl = *reinterpret_cast<const bf8 *>(tb + kBfLoPlane + jj * Ppad * 64 + bfoff);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const bf8 ah = __builtin_bit_cast(bf8, ka[sl][q]), al = __builtin_bit_cast(bf8, kl[sl][q]);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc[q], 0, 0, 0);  // small terms first
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc[q], 0, 0, 0);
                    acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc[q], 0, 0, 0);
                }
            } else {
                const float *tf = reinterpret_cast<const float *>(tb);
                float b[4];
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const int kbl = 4 * jj + e;
                    b[e] = tf[(kbl * NM + mt) * 64 + boff[kbl & 7]];
                }
#pragma unroll
                for (int e = 0; e < 4; e++)
#pragma unroll
                    for (int q = 0; q < NQ; q++)
                        // (the element through a value: a bit_cast of a vector element
                        // expression reads element 0 in this compiler)
                        acc[q] = __builtin_amdgcn_mfma_f32_16x16x4f32(__uint_as_float(ka[sl][q][e]), b[e], acc[q], 0,
                                                                      0, 0);
            }
            const int so = min(cm * UPC + jj + KD, units - 1) * unit_bytes;
#pragma unroll
            for (int q = 0; q < NQ; q++) {
                ka[sl][q] = buffer_load4(kr[q], lo_, so);
                if (BF) kl[sl][q] = buffer_load4(kr[q], lo_ + 16u, so);
            }
            __builtin_amdgcn_sched_barrier(0);  // issue the refill here, a ring ahead of its use
        }
        st.step(c, sh, lds, wave, lane);
        step_barrier();
    }
    for (int c = min(4, nsteps) + nwork; c < nsteps; c++) {
        st.step(c, sh, lds, wave, lane);
        step_barrier();
    }
    __syncthreads();
    // dE[t][m] into LDS (the chunk buffers are free now): lane l holds rows
    // 4 (l/16) + v of the tile, column l%16
    if (chains) {
        const int m = mt * 16 + (lane & 15);
#pragma unroll
        for (int q = 0; q < NQ; q++)
            if (tv[q] && m < d.P)
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int t = tq[q] * 16 + 4 * (lane >> 4) + v;
                    if (t < d.T) lds[t * Ppad + m] = static_cast<double>(acc[q][v]);
                }
    }
}

// Chain waves of the detailed-balance (bdsf) kernel: the trials' model ISF
// as the reference's detailed-balance build sums it (its loop inlines
// matrix_multiply_LxM_by_MxN, deac.cpp:110-118, 1389-1390): per (tau,
// member) a fold of rounded products in position order, the last position
// fused when the genome length is odd.  Register tiles of the outer product
// K[t][k] S[m][k] per position: chain wave w takes members 4 NM w ..
// 4 NM w + 4 NM - 1; lane l takes RL consecutive timeslices of row block
// l % (64 / NM) and the 4 consecutive members of slot l / (64 / NM), RL x 4
// accumulators.  Per position a lane loads its RL values of K (kfold
// [Lpad][fold_rows], timeslices contiguous, D positions ahead, through a
// buffer resource whose wave-uniform offset carries the position) and its
// four trial values (two 16-byte LDS reads, the same address for every lane
// of a slot).  XR: the timeslices past the row blocks, R .. T - 1 (at most
// 4): wave w folds timeslice R + w, lane m member m, K loaded at one address
// by all lanes.
template <int RL, int XR, int NM>
__device__ __forceinline__ void fold_chains(const BatchDesc &d, double *lds, bool chains, int nchunk, int nsteps,
                                            int wave, int lane)
{
    constexpr int C = 128 / NM, Ppad = 16 * NM;
    constexpr int RQ = 64 / NM;               // row blocks per wave (lanes per member slot)
    constexpr int R = RQ * RL;                // timeslices of the row blocks
    constexpr int D = RL <= 2 ? 8 : 16 / RL;  // K prefetch distance in positions (divides 8)
    constexpr int NB = (RL + 1) / 2;          // 16-byte K loads per position
    const int T = d.T, L = d.L, P = d.P, tp = d.fold_rows;
    const int rq = lane % RQ, ms = lane / RQ;
    const int mb = wave * 4 * NM + 4 * ms;   // the lane's first member
    double acc[RL][4];
#pragma unroll
    for (int i = 0; i < RL; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = 0.0;
    const rsrc_t rk = make_rsrc(d.kfold, d.Lpad * tp * 8);  // positions past Lpad read zeros
    const uint32_t vo = rq * RL * 8;
    const int pstride = tp * 8;               // bytes per position
    // the extra timeslice R + wave: its K value at byte offset xo of a
    // position, the same address in every lane
    const int xo = (R + wave) * 8;
    auto load_x = [&](int pos) {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rk, 0u, pos * pstride + xo, 0));
    };
    const bool xv = XR && R + wave < T && lane < P;
    const int xm = lane < Ppad ? lane : 0;
    double accx = 0.0;
    const bool lodd = (L & 1) != 0;
    const int nwork = chains ? nchunk : 0;
    // K of positions p .. p + D - 1 in flight (slot p % D)
    double kb[D][RL], xb[D];
    auto load_k = [&](double (&kv)[RL], int pos) {
        if constexpr (RL == 1) {
            kv[0] = __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rk, vo, pos * pstride, 0));
        } else {
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const double2 v = buffer_load2(rk, vo + 16 * b, pos * pstride);
                kv[2 * b] = v.x;
                kv[2 * b + 1] = v.y;
            }
        }
    };
    if (nwork > 0) {
#pragma unroll
        for (int u = 0; u < D; u++) {
            load_k(kb[u], u);
            if constexpr (XR) xb[u] = load_x(u);
        }
    }
    for (int c = 0; c < min(4, nsteps); c++) step_barrier();
    for (int cm = 0; cm < nwork; cm++) {
        const double *tb = lds + kOffTrial + (cm & 1) * kCD;
        const int k0 = cm * C, n = min(C, L - k0);
        if (n == C && !(lodd && k0 + C == L)) {
#pragma unroll
            for (int s0 = 0; s0 < C; s0 += 8) {
                // the slot's four members sit in one aligned quad of each
                // position's 16-member group: member mb + j at quad entry
                // j ^ ((sl >> 1) & 3); the quad at (mb & 15) ^ ((sl >> 1) & 12),
                // fixed across the 8 positions.  The extra timeslice's member
                // xm at entry (xm & 15) ^ ((sl >> 1) & 15) of its group.
                const double *qb = tb + (s0 >> 2) * NM * 64 + (mb >> 4) * 64 + ((mb & 15) ^ ((s0 >> 1) & 12));
                const double *xq = tb + (s0 >> 2) * NM * 64 + (xm >> 4) * 64;
                const int xs0 = (xm & 15) ^ ((s0 >> 1) & 12);
                auto quad = [&](int u, double2 &a, double2 &b) {
                    const double2 *q = reinterpret_cast<const double2 *>(qb + (u >> 2) * NM * 64 + (u & 3) * 16);
                    a = q[0];
                    b = q[1];
                };
                auto xval = [&](int u) { return xq[(u >> 2) * NM * 64 + (u & 3) * 16 + (xs0 ^ ((u >> 1) & 3))]; };
                // the trial values of position u + 1 are read while position
                // u's products run
                double2 na, nb;
                double nx = 0.0;
                quad(0, na, nb);
                if constexpr (XR) nx = xval(0);
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int sl = s0 + u;
                    const double2 a = na, b = nb;
                    const double xs = nx;
                    if (u < 7) {
                        quad(u + 1, na, nb);
                        if constexpr (XR) nx = xval(u + 1);
                    }
                    const double v[4] = {a.x, a.y, b.x, b.y};
                    const int sw = (u >> 1) & 3;
                    double kc[RL];
#pragma unroll
                    for (int i = 0; i < RL; i++) kc[i] = kb[u % D][i];
#pragma unroll
                    for (int i = 0; i < RL; i++)
#pragma unroll
                        for (int j = 0; j < 4; j++) acc[i][j] = acc[i][j] + v[j ^ sw] * kc[i];
                    if constexpr (XR) accx = accx + xs * xb[u % D];
                    __builtin_amdgcn_sched_barrier(0);
                    const int pn = k0 + sl + D;
                    load_k(kb[u % D], pn);
                    if constexpr (XR) xb[u % D] = load_x(pn);
                    // keep each position's loads where they are, after its
                    // products: the scheduler would otherwise sink the K
                    // loads to the end of the 8 positions, or start one
                    // before the products that read the slot's old values,
                    // and wait for it there
                    __builtin_amdgcn_sched_barrier(0);
                }
            }
        } else {
            // the last chunk: partial, or holding the fused last position
            for (int sl = 0; sl < n; sl++) {
                const int k = k0 + sl;
                const bool f = lodd && k == L - 1;
                double v[4], kv[RL];
#pragma unroll
                for (int j = 0; j < 4; j++) v[j] = tb[trial_index<NM>(mb + j, sl)];
                load_k(kv, k);
#pragma unroll
                for (int i = 0; i < RL; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = f ? fma(v[j], kv[i], acc[i][j]) : acc[i][j] + v[j] * kv[i];
                if constexpr (XR) {
                    const double xs = tb[trial_index<NM>(xm, sl)], xk = load_x(k);
                    accx = f ? fma(xs, xk, accx) : accx + xs * xk;
                }
            }
        }
        step_barrier();
    }
    for (int c = min(4, nsteps) + nwork; c < nsteps; c++) step_barrier();
    __syncthreads();
    // E[t][m] into LDS (the chunk buffers are free now)
    if (chains) {
#pragma unroll
        for (int i = 0; i < RL; i++) {
            const int t = rq * RL + i;
#pragma unroll
            for (int j = 0; j < 4; j++)
                if (t < T && mb + j < P) lds[t * Ppad + mb + j] = acc[i][j];
        }
        if (xv) lds[(R + wave) * Ppad + lane] = accx;
    }
}

// One pass over the genome in chunks.  kind: kPassInit (model ISF of the
// stored population), kPassEvolve (generation g), kPassApply (write back the
// genomes accepted in generation g), kPassPre (masks only), kPassNorm
// (per-generation renormalization, FP64 kernel: the write-back of generation
// g-1's acceptances and the zeroth moment of each trial of generation g,
// before its kPassEvolve, which then forms the trials again, scaled, and
// rebuilds nothing), kPassNone (no steps: the mask pass of a resumed
// launch).  mgen: the generation whose masks the workers draw (-1: none);
// wgen: the generation the walk wave walks (-1: none).
template <int PREC, int NT, int NM, int REM>
__device__ __forceinline__ void run_pass(const BatchDesc &d, BatchShared &sh, double *lds, int kind, long g, long mgen,
                                         long wgen)
{
    constexpr int C = 128 / NM, Ppad = 16 * NM;
    // kernel-tile prefetch depth in pairs of 4-position blocks; it must
    // divide C/8 (4 block pairs per chunk at NM = 4), since the chain loop
    // keeps block pair g in slot (g mod C/8) mod KD
    constexpr int KD = std::min(NT == 1 ? (REM ? 4 : 8) : (NT == 2 ? 4 : 2), C / 8);
    static_assert((C / 8) % KD == 0, "the K prefetch depth must divide the block pairs of a chunk");
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    const int nchunk = d.nchunk;
    constexpr bool F64 = PREC == kPrecFp64 || PREC == kPrecFold;  // FP64 trial values
    // mixed K.S instantiations with REM = 1 take moment terms (first, third, -1st)
    constexpr bool MOM = (PREC == kPrecFp32 || PREC == kPrecBf16) && REM != 0;
    static_assert(!MOM || (C / 2) * Ppad * 2 <= kPartDoubles, "the partial moment sums fit their region");
    static_assert(kMixedIsf <= kOffPart, "the mixed model ISF stays below the partial sums");
    const bool chains = kind == kPassInit || kind == kPassEvolve;
    const int nsteps = kind == kPassApply ? nchunk + 2 : kind == kPassNone ? 0 : nchunk + 4;

    if (wave < kWorker0) {
        if constexpr (PREC == kPrecFold) {
            fold_chains<NT / 2, NT % 2, NM>(d, lds, chains, nchunk, nsteps, wave, lane);  // NT: fold shape 2 RL + XR
        } else if constexpr (PREC != kPrecFp64) {
            mixed_chains<PREC, NT, NM>(d, sh, lds, kind, g, chains, nchunk, nsteps, wave, lane);
        } else {
            // ------------------------------------------------ chain waves
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
            int boff[8];
            {
                const int q = lane >> 4, j = lane & 15;
#pragma unroll
                for (int u = 0; u < 8; u++) boff[u] = q * 16 + (j ^ ((2 * u + (q >> 1)) & 15));
            }
            d4 acc[NT][NM];
#pragma unroll
            for (int i = 0; i < NT; i++)
#pragma unroll
                for (int mt = 0; mt < NM; mt++) acc[i][mt] = d4{0.0, 0.0, 0.0, 0.0};
            // K tiles through buffer resources: the lane's byte offset is fixed
            // and the block offset wave-uniform (a scalar register), so the
            // refills take no vector address arithmetic.
            rsrc_t ktr[NT];
#pragma unroll
            for (int i = 0; i < NT; i++)
                ktr[i] = make_rsrc(d.ktile + static_cast<size_t>(tload[i]) * nkb2 * (2 * kWave), nkb2 * (2 * kWave) * 8);
            const uint32_t lob = 16u * lane;
            // Rows past the last full tile (REM): one 4x4x4 MFMA per 4 positions
            // for member tile rmt = wave % NM, on waves 0 .. NM-1.  The nrem
            // remainder rows are stored nrem x 4 values per 4-position block,
            // lane l reading row l & 3 at position l >> 4; lanes of rows past
            // nrem read past the buffer's end, i.e. zeros, into chains whose
            // results are discarded.
            const int nr = max(d.nrem, 1);
            const rsrc_t krm = make_rsrc(d.krem, nkb2 * nr * 64);
            const int rmt = wave % NM;
            const uint32_t lrob = (lane & 3) < nr ? ((lane & 3) + nr * (lane >> 4)) * 16u : 0x40000000u;
            double accr = 0.0;
            const int nwork = chains ? nchunk : 0;
            // One copy of the loop per remainder member tile RT (the 4x4x4 MFMA
            // takes its B register directly), and one without the remainder for
            // the other waves.  Each copy issues its own K prefetch before the
            // four pipeline-filling steps; steps 4 .. nchunk+3 run the chains of
            // chunk c-4.  The hot loop has no branch, so the waits for the K
            // prefetch are exact across its iterations.
            auto chain_loop = [&](auto rt, auto rem) {
                constexpr int RT = decltype(rt)::value;
                constexpr bool RM = decltype(rem)::value;
                double2 kr[KD][NT];
                double2 krr[KD];
#pragma unroll
                for (int p = 0; p < KD; p++)
#pragma unroll
                    for (int i = 0; i < NT; i++) kr[p][i] = buffer_load2(ktr[i], lob, min(p, nkb2 - 1) * (2 * kWave * 8));
                if (RM) {
#pragma unroll
                    for (int p = 0; p < KD; p++) krr[p] = buffer_load2(krm, lrob, min(p, nkb2 - 1) * (nr * 64));
                }
                for (int c = 0; c < min(4, nsteps); c++) step_barrier();
                for (int cm = 0; cm < nwork; cm++) {
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
                                for (int mt = 0; mt < NM; mt++) bn[e][mt] = tb[(kbl * NM + mt) * kWave + boff[kbl & 7]];
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
                            if (RM) {
                                const double ar = e ? krr[sl].y : krr[sl].x;
                                accr = __builtin_amdgcn_mfma_f64_4x4x4f64(ar, bc[e][RT], accr, 0, 0, 0);
                            }
                        }
#pragma unroll
                        for (int e = 0; e < 2; e++)
#pragma unroll
                            for (int mt = 0; mt < NM; mt++) bc[e][mt] = bn[e][mt];
                        const int gn = min(g2 + KD, nkb2 - 1);
#pragma unroll
                        for (int i = 0; i < NT; i++) kr[sl][i] = buffer_load2(ktr[i], lob, gn * (2 * kWave * 8));
                        if (RM) krr[sl] = buffer_load2(krm, lrob, gn * (nr * 64));
                        __builtin_amdgcn_sched_barrier(0);  // issue the refill here, a full step ahead of its use
                    }
                    step_barrier();
                }
                for (int c = min(4, nsteps) + nwork; c < nsteps; c++) step_barrier();
            };
            if (!REM || wave >= NM) chain_loop(std::integral_constant<int, 0>{}, std::false_type{});
            else if (rmt == 0) chain_loop(std::integral_constant<int, 0>{}, std::true_type{});
            else if (NM < 4 || rmt == NM - 1) chain_loop(std::integral_constant<int, NM - 1>{}, std::true_type{});
            else if (rmt == 1) chain_loop(std::integral_constant<int, NM == 4 ? 1 : 0>{}, std::true_type{});
            else chain_loop(std::integral_constant<int, NM == 4 ? 2 : 0>{}, std::true_type{});
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
        }
    } else if (wave < kWalkWave) {
        // ------------------------------------------------ workers
        // Worker wk owns rows wk and wk + 8 of the chunk buffers ([Ppad][C]
        // doubles, 16 rows of 128).  Lane l of row R holds positions p, p + 1
        // (p = 2l mod C) of member m = (128 R + 2l) / C in every stage: the
        // global load of the chunk, the rebuild of the genomes accepted in
        // the previous generation, and this generation's trial vectors.
        const int wk = wave - kWorker0;
        // Descriptor fields held in scalar registers for the whole pass.
        const int nP = __builtin_amdgcn_readfirstlane(d.P), nG = __builtin_amdgcn_readfirstlane(d.G);
        const int nL = __builtin_amdgcn_readfirstlane(d.L), nLpad = __builtin_amdgcn_readfirstlane(d.Lpad);
        const int n2 = __builtin_amdgcn_readfirstlane(d.nsign) == 2 ? 1 : 0;
        const int nLw = __builtin_amdgcn_readfirstlane(d.Lw);
        double *const popp = uniform_ptr(d.pop);
        uint32_t *const maskp = uniform_ptr(d.masks);
        // fabs of a DE value, or not (allow_negative): an and of the high word
        const uint32_t absmask = __builtin_amdgcn_readfirstlane(d.allow_neg) ? 0xffffffffu : 0x7fffffffu;
        constexpr int kWpc = C / 32;  // mask words per member and chunk
        const bool renorm = (F64 || MOM) && d.renorm;
        const bool trials = kind == kPassEvolve || kind == kPassNorm;
        const bool rebuild = kind == kPassApply || kind == kPassNorm || (kind == kPassEvolve && !renorm);
        const long pg = trials ? g - 1 : g;                   // generation whose accepted trials are rebuilt
        const int pslot = static_cast<int>((pg + 3) % 3);
        const int pprev = static_cast<int>((pg + 2) % 3);     // generation pg-1
        const int gslot = static_cast<int>((g + 3) % 3);
        const int aslot = static_cast<int>((g + 2) % 3);      // generation g-1
        const int p = (2 * lane) % C;
        // Mixed K.S with moment constraints (MOM): the first and third
        // moments of each trial, the lane's two positions per chunk summed in
        // FP64 across the pass (the mixed fitness is an approximation: the
        // FP64 evaluation after the slice recomputes every reported fitness
        // in the reference's order); the weights of the trial chunk are
        // loaded at the top of its step.
        const bool wmom = MOM && kind == kPassEvolve && (d.use_first || d.use_third);
        const bool wnorm = MOM && kind == kPassNorm;  // the trials' zeroth moments (renormalization)
        const gcdouble *const mtp = (const gcdouble *)uniform_ptr(d.mterm);
        double pm1[kT], pm3[kT];
        double2 wm1 = make_double2(0.0, 0.0), wm3 = make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < kT; u++) pm1[u] = pm3[u] = 0.0;
        // Per-row constants of the pass.  Padded members (m >= P) compute on
        // the last member's data: their trial columns only feed their own
        // (discarded) model values.
        // Offset (doubles) of the lane's pair in a chunk buffer, row 0; row u
        // adds 1024.  Member of row u, and the member whose data it uses.
        const int own = wk * 128 + 2 * lane;
        auto mrow = [&](int u) { return (own + u * 1024) / C; };
        auto fm = [&](int u) { return min(mrow(u), nP - 1); };
        // trial-buffer index of position p of row 0 (the others follow from it)
        const int td0 = trial_index<NM>(mrow(0), p);
        int od[kT][3], odb[kT][3], mo[kT], mob[kT];
        uint32_t poff[kT];
        uint64_t bm[kT];  // lanes whose member was accepted in generation pg
#pragma unroll
        for (int u = 0; u < kT; u++) {
            const int m = mrow(u), cm = fm(u);
            const int cq = cm / d.Pi;  // instance of the row's member, its row in that instance
            poff[u] = d.pop_row0[cq] + static_cast<uint32_t>((cm - cq * d.Pi) * nLpad + p) * 8u;
            od[u][0] = od[u][1] = od[u][2] = m * C + p;
            odb[u][0] = odb[u][1] = odb[u][2] = m * C + p;
            mo[u] = mob[u] = 0;
            if (trials) {
                const WalkRec &r = sh.rec[g & 3][cm];
#pragma unroll
                for (int k = 0; k < 3; k++) od[u][k] = static_cast<int>(r.mutant[k] & (kBatchMaxP - 1)) * C + p;
                mo[u] = (cm * 2 + (sh.accepted[aslot][cm] & 1)) * kWpc + (p >> 5);
            }
            // The walk record of pg is read only for accepted members, and
            // only for generations evolved in this launch: the record of the
            // generation before a launch's first one was never written in it
            // (a resumed launch holds that generation's acceptances, for the
            // crossover candidates, but its genomes were written back).
            const bool acc = rebuild && pg >= d.gen0 && m < nP && sh.accepted[pslot][cm];
            if (acc) {
                const WalkRec &r = sh.rec[pg & 3][cm];
#pragma unroll
                for (int k = 0; k < 3; k++) odb[u][k] = static_cast<int>(r.mutant[k] & (kBatchMaxP - 1)) * C + p;
                mob[u] = Ppad * 2 * kWpc + (cm * 2 + (sh.accepted[pprev][cm] & 1)) * kWpc + (p >> 5);
            }
            bm[u] = __ballot(acc);
        }
        // Crossover words of chunk c, loaded by workers 0-3: this
        // generation's (both candidates) and generation pg's.
        uint32_t moff = 0;
        if (wk < kMaskWords / kWave) {
            const int e = wk * kWave + lane;
            const int w = e % kWpc, cand = (e / kWpc) & 1, m = (e / (2 * kWpc)) % Ppad;
            const int slot = e / (2 * kWpc * Ppad) == 0 ? gslot : pslot;
            moff = static_cast<uint32_t>(((slot * Ppad + m) * 2 + cand) * nLw + w) * 4u;
        }

        // DE values at positions p, p+1 of chunk cc for one row: operands
        // (parent, r1, r2, r3) from chunk buffer buf, crossover word mr[wi],
        // jDE weight F by frequency sign (deac.cpp:1261-1311).  Chunks
        // holding position 0, position G (the zero frequency of the negative
        // half, which repeats position 0: deac.cpp:1305-1308) or padding take
        // the per-position path; elsewhere the sign is uniform and one F
        // serves the chunk.  Operands of both rows are read before either is
        // computed (one LDS round trip per stage).
        struct Ops {
            double2 x0, x1, x2, x3;
            double f0, f1;  // F (fast path); F+, F- (per-position path)
            uint32_t w;
        };
        auto load_ops = [&](const double *buf, const uint32_t *mr, int wi, const int (&o)[3], int u) -> Ops {
            Ops a;
            a.x0 = *reinterpret_cast<const double2 *>(buf + own + u * 1024);
            a.x1 = *reinterpret_cast<const double2 *>(buf + o[0]);
            a.x2 = *reinterpret_cast<const double2 *>(buf + o[1]);
            a.x3 = *reinterpret_cast<const double2 *>(buf + o[2]);
            a.w = mr[wi] >> (p & 31);
            return a;
        };
        auto de_fast = [&](const Ops &a) -> double2 {
            double2 v;
            v.x = (a.w & 1) ? de_abs(fma(a.f0, a.x2.x - a.x3.x, a.x1.x), absmask) : a.x0.x;
            v.y = (a.w & 2) ? de_abs(fma(a.f0, a.x2.y - a.x3.y, a.x1.y), absmask) : a.x0.y;
            return v;
        };
        auto de_slow = [&](const Ops &a, const double *buf, const uint32_t *mr, int wi, const int (&o)[3], int u,
                           int cc, double *y0) -> double2 {
            const int m = mrow(u);
            double r[2];
#pragma unroll
            for (int e = 0; e < 2; e++) {
                const int sp = cc * C + p + e;
                const double F = (n2 && sp >= nG) ? a.f1 : a.f0;
                const double a0 = e ? a.x0.y : a.x0.x, a1 = e ? a.x1.y : a.x1.x, a2 = e ? a.x2.y : a.x2.x,
                             a3 = e ? a.x3.y : a.x3.x;
                double t = ((a.w >> e) & 1) ? de_abs(fma(F, a2 - a3, a1), absmask) : a0;
                if (n2 && sp == nG) {
                    if (cc == 0) {
                        // position 0 of the same chunk, positive frequencies
                        const uint32_t w0 = mr[wi - (p >> 5)];
                        t = (w0 & 1) ? de_abs(fma(a.f0, buf[o[1] - p] - buf[o[2] - p], buf[o[0] - p]), absmask)
                                     : buf[m * C];
                    } else {
                        t = y0[m];
                    }
                }
                if (sp == 0) y0[m] = t;
                if (sp >= nL) t = 0.0;
                r[e] = t;
            }
            double2 v;
            v.x = r[0];
            v.y = r[1];
            return v;
        };
        // One DE stage over both rows: chunk cc of buffer buf, crossover words
        // mr (+ wi), donors o, weights of generation gen, position-0 values y0.
        // delta: return trial - parent (the mixed kernel's dS) instead of the trial.
        auto de_stage = [&](const double *buf, const uint32_t *mr, const int (&wi)[kT], const int (&o)[kT][3], long gen,
                            int cc, double *y0, double2 (&v)[kT], bool delta) {
            Ops a[kT];
            if (!(cc == 0 || (n2 && cc == nG / C) || (cc + 1) * C > nL)) {
                const int neg = (n2 && cc * C >= nG) ? 1 : 0;
#pragma unroll
                for (int u = 0; u < kT; u++) {
                    a[u] = load_ops(buf, mr, wi[u], o[u], u);
                    a[u].f0 = sh.f_trial[gen & 1][fm(u)][neg];
                }
#pragma unroll
                for (int u = 0; u < kT; u++) v[u] = de_fast(a[u]);
            } else {
#pragma unroll
                for (int u = 0; u < kT; u++) {
                    a[u] = load_ops(buf, mr, wi[u], o[u], u);
                    a[u].f0 = sh.f_trial[gen & 1][fm(u)][0];
                    a[u].f1 = sh.f_trial[gen & 1][fm(u)][1];
                }
#pragma unroll
                for (int u = 0; u < kT; u++) v[u] = de_slow(a[u], buf, mr, wi[u], o[u], u, cc, y0);
            }
            if (delta) {
                // exact zero where the crossover keeps the parent
#pragma unroll
                for (int u = 0; u < kT; u++) {
                    v[u].x -= a[u].x0.x;
                    v[u].y -= a[u].x0.y;
                }
            }
        };

        // load(c) goes to registers at the top of step c and into LDS at its
        // bottom: plain loads, so no LDS access in between waits for them (an
        // LDS DMA would make every LDS read wait for the copy in flight).
        // rebuild(c-1) computes the genomes accepted in pg from the chunk as
        // loaded; they replace it in place (LDS and HBM) at the top of the
        // next step, once every rebuild has read the old values; trial(c-3)
        // then reads this generation's genomes from one buffer.
        double2 pv[kT], rb[kT];
        uint32_t mw = 0;
        // byte offsets of the lane's pair in the BF16X3 trial buffer
        int bfo[kT];
#pragma unroll
        for (int u = 0; u < kT; u++) bfo[u] = bf_trial_offset<NM>(mrow(u), p);
        for (int c = 0; c < nsteps; c++) {
            // (the mixed kernel's chain waves stage the chunks instead: Stager;
            // in the FP64 kernel their registers hold the kernel-tile ring)
            const bool ld = F64 && kind != kPassPre && c < nchunk;
            if constexpr (MOM) {
                if ((wmom || wnorm) && c >= 3 && c - 3 < nchunk) {
                    const size_t k = static_cast<size_t>(c - 3) * C + p + (wnorm ? 2 * static_cast<size_t>(nLpad) : 0);
                    wm1 = make_double2(mtp[k], mtp[k + 1]);
                    if (wmom) wm3 = make_double2(mtp[nLpad + k], mtp[nLpad + k + 1]);
                }
            }
            if (ld) {
                const gchar *pb = (const gchar *)popp + static_cast<size_t>(c) * C * 8;
#pragma unroll
                for (int u = 0; u < kT; u++) pv[u] = load2_stream((const gcdouble *)(pb + poff[u]));
                if (wk < kMaskWords / kWave)
                    mw = *(const guint *)((const gchar *)maskp + static_cast<size_t>(c) * kWpc * 4 + moff);
            }
            // rebuild write-back of chunk c-2
            const int cw = c - 2;
            if (rebuild && cw >= 0 && cw < nchunk) {
                double *buf = lds + kOffPop + (cw % 4) * kCD;
                gchar *gb = (gchar *)popp + static_cast<size_t>(cw) * C * 8;
#pragma unroll
                for (int u = 0; u < kT; u++)
                    if (bm[u] && ((bm[u] >> lane) & 1)) {
                        *reinterpret_cast<double2 *>(buf + own + u * 1024) = rb[u];
                        gdouble *q = (gdouble *)(gb + poff[u]);
                        __builtin_nontemporal_store(rb[u].x, q);
                        __builtin_nontemporal_store(rb[u].y, q + 1);
                    }
            }
            // trial of chunk c-3 (or, for the initial evaluation, the stored
            // genomes) into the trial buffer in MFMA B order
            const int ct = c - 3;
            if ((chains || kind == kPassNorm) && ct >= 0 && ct < nchunk) {
                const double *buf = lds + kOffPop + (ct % 4) * kCD;
                double *tb = lds + kOffTrial + (ct & 1) * kCD;
                double2 v[kT];
                if (kind == kPassInit) {
                    // (FP64 only: the mixed kernel is never given a population to evaluate)
#pragma unroll
                    for (int u = 0; u < kT; u++) {
                        const double2 x = *reinterpret_cast<const double2 *>(buf + own + u * 1024);
                        const int sp = ct * C + p;
                        v[u].x = sp < nL ? x.x : 0.0;
                        v[u].y = sp + 1 < nL ? x.y : 0.0;
                    }
                } else {
                    de_stage(buf, sh.mring[ct % 4], mo, od, g, ct, sh.y0t, v, !F64 && !MOM);
                    if (renorm && kind == kPassEvolve) {
                        // the trial times zeroth_moment / norm (deac.cpp:1353-1360)
#pragma unroll
                        for (int u = 0; u < kT; u++) {
                            const double sc = sh.rscale[g & 1][fm(u)];
                            v[u].x *= sc;
                            v[u].y *= sc;
                        }
                    }
                    if constexpr (MOM) {
                        // the trial's moment terms (or, in the renormalization
                        // pass, its zeroth moment), then dS = trial - parent
#pragma unroll
                        for (int u = 0; u < kT; u++) {
                            if (wmom || wnorm) pm1[u] += v[u].x * wm1.x + v[u].y * wm1.y;
                            if (wmom) pm3[u] += v[u].x * wm3.x + v[u].y * wm3.y;
                            const double2 x0 = *reinterpret_cast<const double2 *>(buf + own + u * 1024);
                            v[u].x -= x0.x;
                            v[u].y -= x0.y;
                        }
                    }
                }
                if constexpr (F64) {
#pragma unroll
                    for (int u = 0; u < kT; u++) {
                        const int t0 = NM >= 2 ? td0 + 32 * NM * u : td0 ^ (8 * u);  // position p
                        tb[t0] = v[u].x;
                        tb[t0 + 16] = v[u].y;                                     // position p + 1
                    }
                } else if constexpr (PREC == kPrecFp32) {
                    // dS = trial - parent in FP64, rounded once to FP32
                    float *tf = reinterpret_cast<float *>(tb);
#pragma unroll
                    for (int u = 0; u < kT; u++) {
                        const int t0 = NM >= 2 ? td0 + 32 * NM * u : td0 ^ (8 * u);
                        tf[t0] = static_cast<float>(v[u].x);
                        tf[t0 + 16] = static_cast<float>(v[u].y);
                    }
                } else {
                    // dS = hi + lo: hi = bf16(dS), lo = bf16(dS - hi) in FP32
                    char *tc = reinterpret_cast<char *>(tb);
#pragma unroll
                    for (int u = 0; u < kT; u++) {
                        const float a = static_cast<float>(v[u].x);
                        const float b = static_cast<float>(v[u].y);
                        const uint32_t h = pack_bf16(a, b);
                        const uint32_t l = pack_bf16(a - __uint_as_float(h << 16), b - __uint_as_float(h & 0xffff0000u));
                        *reinterpret_cast<uint32_t *>(tc + bfo[u]) = h;
                        *reinterpret_cast<uint32_t *>(tc + kBfLoPlane + bfo[u]) = l;
                    }
                }
            }
            // rebuild of chunk c-1 (generation pg), both rows when either
            // holds an accepted member
            const int cb = c - 1;
            if (rebuild && cb >= 0 && cb < nchunk && (bm[0] | bm[kT - 1])) {
                de_stage(lds + kOffPop + (cb % 4) * kCD, sh.mring[cb % 4], mob, odb, pg, cb, sh.y0r, rb, false);
                if (renorm) {
#pragma unroll
                    for (int u = 0; u < kT; u++) {
                        const double sc = sh.rscale[pg & 1][fm(u)];
                        rb[u].x *= sc;
                        rb[u].y *= sc;
                    }
                }
            }
            // load(c) lands
            if (ld) {
#pragma unroll
                for (int u = 0; u < kT; u++)
                    *reinterpret_cast<double2 *>(lds + kOffPop + (c % 4) * kCD + own + u * 1024) = pv[u];
                if (wk < kMaskWords / kWave) sh.mring[c % 4][wk * kWave + lane] = mw;
            }
            step_barrier();
        }
        __syncthreads();
        if constexpr (MOM) {
            if (wmom || wnorm) {
#pragma unroll
                for (int u = 0; u < kT; u++)
                    if (mrow(u) < nP) {
                        double *q = lds + kOffPart + (mrow(u) * (C / 2) + p / 2) * 2;
                        q[0] = pm1[u];
                        q[1] = pm3[u];
                    }
            }
        }
    } else if (wave >= kRngWave0) {
        // ------------------------------------------------ mask waves
        // The next generation's crossover draws, Lsub per lane per pass.
        // Draws are spread evenly over the steps (a cumulative target, in
        // groups of 8): these waves share their SIMDs with the chains, whose
        // MFMAs wait for every VALU instruction issued beside them.
        MaskGen mg;
        const bool draws = mgen >= 0;
        const int lsub = d.Lsub;
        if (draws) mask_setup<NM>(mg, d, sh, mgen, kind == kPassEvolve, g, (wave - kRngWave0) * kWave + lane);
        int drawn = 0;
        for (int c = 0; c < nsteps; c++) {
            if (draws) {
                const int target = ((c + 1) * lsub + nsteps - 1) / nsteps;
                const int q = (target + 7) / 8 * 8 - drawn;
                if (q > 0) {
                    mask_draws(mg, q, d.Lw);
                    drawn += q;
                }
            }
            step_barrier();
        }
        DEAC_WAIT_VM(0);
        __syncthreads();
    } else {
        // ------------------------------------------------ walk wave
        // It also folds the moment rows of the pass's genomes (FP64 kernel,
        // moment constraints): lane m reads member m's values of chunk c - 4
        // from the trial buffer in the step the chains contract it, and adds
        // the products in the reference's order (deac.cpp:102-108 as the
        // reference build inlines it, deac.cpp:1420-1477): rounded products
        // one at a time, the last position of each sign fused when G is odd.
        // The streams of the workgroup's instances one after another: items
        // 2 Pi q .. 2 Pi (q + 1) - 1 are instance q's.
        const int items = 2 * d.P, item_q = 2 * d.Pi;
        const int per = wgen >= 0 ? (items + nsteps - 1) / nsteps : 0;
        int cq = 0;
        uint64_t ws[4] = {sh.walk_state[0][0], sh.walk_state[0][1], sh.walk_state[0][2], sh.walk_state[0][3]};
        int next = 0;
        // Moment rows (first, third) or the normalization row (kPassNorm):
        // the weights of chunk c are the same for every member; lane l loads
        // positions l, l + 64 of the chunk a step ahead, and the fold reads
        // them back with v_readlane.
        const bool moments = F64 && chains && (d.use_first || d.use_third);
        const bool norms = F64 && kind == kPassNorm;
        const bool folds = moments || norms;
        const gcdouble *const mt = (const gcdouble *)uniform_ptr(d.mterm);
        const int nL = __builtin_amdgcn_readfirstlane(d.L), nG = __builtin_amdgcn_readfirstlane(d.G);
        const int nLpad = __builtin_amdgcn_readfirstlane(d.Lpad);
        const bool godd = (nG & 1) != 0;
        const int row0 = norms ? 2 : 0;  // the weight rows folded: row0 (and row0 + 1 for the moments)
        constexpr int WPL = (C + kWave - 1) / kWave;
        double wn[2][WPL], wc[2][WPL];
        auto load_w = [&](int chunk) {
#pragma unroll
            for (int r = 0; r < 2; r++)
#pragma unroll
                for (int j = 0; j < WPL; j++) {
                    const int sl = j * kWave + lane, k = chunk * C + sl;
                    wn[r][j] = (sl < C && k < nL && (r == 0 || moments)) ? mt[(row0 + r) * nLpad + k] : 0.0;
                }
        };
        if (folds && nchunk > 0) load_w(0);
        double m1 = 0.0, m3 = 0.0;
        for (int c = 0; c < nsteps; c++) {
            const int cm = c - 4;
            if (folds && cm >= 0 && cm < nchunk) {
#pragma unroll
                for (int r = 0; r < 2; r++)
#pragma unroll
                    for (int j = 0; j < WPL; j++) wc[r][j] = wn[r][j];
                if (cm + 1 < nchunk) load_w(cm + 1);  // lands during this step and the next barrier
                const double *tb = lds + kOffTrial + (cm & 1) * kCD;
                const int k0 = cm * C, n = min(C, nL - k0);
                const int own = lane < d.P ? lane : 0;  // lanes past P fold member 0's values, discarded
                auto weight = [&](int r, int sl) {
                    return readlane_f64(WPL == 1 || sl < kWave ? wc[r][0] : wc[r][WPL - 1], sl & (kWave - 1));
                };
                const bool fused = godd && ((nG - 1 >= k0 && nG - 1 < k0 + n) || (nL - 1 >= k0 && nL - 1 < k0 + n));
                if (n == C && !fused) {
                    // a whole chunk without a fused position: eight trial
                    // values read together, then their products added in order
                    for (int s0 = 0; s0 < C; s0 += 8) {
                        double v[8];
#pragma unroll
                        for (int j = 0; j < 8; j++) v[j] = tb[trial_index<NM>(own, s0 + j)];
                        if (norms) {
#pragma unroll
                            for (int j = 0; j < 8; j++) m1 = m1 + v[j] * weight(0, s0 + j);
                        } else {
#pragma unroll
                            for (int j = 0; j < 8; j++) {
                                m1 = m1 + v[j] * weight(0, s0 + j);
                                m3 = m3 + v[j] * weight(1, s0 + j);
                            }
                        }
                    }
                } else {
                    for (int sl = 0; sl < n; sl++) {
                        const int k = k0 + sl;
                        const double v = tb[trial_index<NM>(own, sl)];
                        const double a = weight(0, sl);
                        const bool f = godd && (k == nG - 1 || k == nL - 1);
                        m1 = f ? fma(v, a, m1) : m1 + v * a;
                        if (!norms) {
                            const double b = weight(1, sl);
                            m3 = f ? fma(v, b, m3) : m3 + v * b;
                        }
                    }
                }
            }
            const int stop = min(items, (c + 1) * per);
            for (; next < stop; next++) {
                const int q = next / item_q;
                if (q != cq) {  // the next instance's stream
                    if (lane == 0)
                        for (int w = 0; w < 4; w++) sh.walk_state[cq][w] = ws[w];
                    for (int w = 0; w < 4; w++) ws[w] = sh.walk_state[q][w];
                    cq = q;
                }
                walk_item(d, sh, ws, wgen, q, next - q * item_q, lane);
            }
            step_barrier();
        }
        if (moments && lane < d.P) {
            sh.mom[lane][0] = m1;
            sh.mom[lane][1] = m3;
        }
        if (norms && lane < d.P) sh.norm[lane] = m1;
        if (lane == 0)
            for (int w = 0; w < 4; w++) sh.walk_state[cq][w] = ws[w];
        __syncthreads();
    }
    __syncthreads();
}

}  // namespace

template <int PREC, int NT, int NM, int REM>
__global__ __launch_bounds__(kBatchThreads) void batch_kernel(const BatchDesc *__restrict__ descs)
{
    constexpr int Ppad = 16 * NM;
    constexpr bool F64 = PREC == kPrecFp64 || PREC == kPrecFold;  // FP64 trial values
    constexpr bool MOM = (PREC == kPrecFp32 || PREC == kPrecBf16) && REM != 0;  // mixed, moment terms
    extern __shared__ __attribute__((aligned(16))) double lds[];  // 16-byte aligned: ds_*_b128 at full rate
    __shared__ BatchShared sh;
    const BatchDesc &d = descs[blockIdx.x];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave), lane = threadIdx.x & (kWave - 1);
    const int P = d.P;
    const gcdouble *obs = (const gcdouble *)d.obs, *sig = (const gcdouble *)d.sig;

#ifdef DEAC_ROLE_TIMERS
    if (threadIdx.x < kBatchWaves) g_wait[threadIdx.x] = 0;
    const unsigned long long tstart = __builtin_amdgcn_s_memtime();
#endif
    for (int k = threadIdx.x; k < 256 * 4; k += kBatchThreads) sh.rows[k] = ((const gcu64 *)d.walk_rows)[k];
    for (int k = threadIdx.x; k < d.T; k += kBatchThreads) {
        sh.obs[k] = obs[k];
        sh.sig[k] = sig[k];
        sh.kscale[k] = F64 ? 1.0 : ((const gcdouble *)d.kscale)[k];
    }
    const bool evaluate = d.mode == 1;  // fitness (and epar) of the stored population only
    for (int k = threadIdx.x; k < 2 * kCD; k += kBatchThreads) lds[kOffTrial + k] = 0.0;  // padded members stay 0
    if (threadIdx.x < kBatchMaxP) {
        const int m = threadIdx.x;
        const int q = m < P ? m / d.Pi : 0, i = m - q * d.Pi;  // instance, its member
        const gcdouble *fit = (const gcdouble *)d.fitness[q], *cr = (const gcdouble *)d.cr_old[q],
                       *fw = (const gcdouble *)d.f_old[q];
        sh.fitness[m] = m < P ? fit[i] : 0.0;
        for (int sg = 0; sg < 2; sg++) {
            sh.cr_old[m][sg] = m < P ? cr[i * 2 + sg] : 0.0;
            sh.f_old[m][sg] = m < P ? fw[i * 2 + sg] : 0.0;
        }
        for (int k = 0; k < 3; k++) sh.accepted[k][m] = 0;
        if (d.resume && !evaluate && m < P)
            sh.accepted[(d.gen0 + 2) % 3][m] = ((const __attribute__((address_space(1))) int *)d.accp[q])[i];
    }
    __syncthreads();

    // Issue priority: the chain and worker waves set the step time; the
    // random-stream waves, which run ahead of their consumers, fill in.
    if (wave < kWalkWave) __builtin_amdgcn_s_setprio(1);

    // Prologue: the walk of the first generation whose masks are drawn.
    const long g0 = d.gen0;
    if (wave == kWalkWave && !evaluate) {
        for (int q = 0; q < d.Q; q++) {
            const gcu64 *rng = (const gcu64 *)d.rng[q];
            uint64_t ws[4] = {rng[0], rng[1], rng[2], rng[3]};
            // a resumed launch walks the next generation too (the mask pass
            // that would walk it is skipped); one loop, so one copy of the walk
            const int items = (d.resume ? 4 : 2) * d.Pi;
            for (int it = 0; it < items; it++) {
                const int gi = it >= 2 * d.Pi ? 1 : 0;
                walk_item(d, sh, ws, g0 + gi, q, it - 2 * d.Pi * gi, lane);
            }
            if (lane == 0)
                for (int w = 0; w < 4; w++) sh.walk_state[q][w] = ws[w];
        }
    }
    __syncthreads();

    // Pass sequence: [initial evaluation | masks of gen0], then one pass per
    // generation, then the write-back of the last generation's acceptances.
    // (The mixed kernel never evaluates from scratch: its initial fitness and
    // epar come from an evaluation launch of the FP64 kernel.)
    int kind = (PREC == kPrecFp64 && d.init) ? kPassInit : kPassPre;
    long g = g0;      // generation at the top of the loop
    long last = -1;   // last generation evolved in this launch
    const int tlast = (d.T & 1) ? d.T - 1 : -1;  // fused element of the chi^2 sum
    gdouble *const epar = (gdouble *)d.epar;     // [T][Ppad] parent model ISF, FP64
    bool normed = false;  // this generation's renormalization pass is done
    for (;;) {
        // per-generation renormalization (deac.cpp:1342-1364): the trials'
        // zeroth moments first, then the pass that scales and contracts them
        // (one call site of the pass, so its code appears once)
        const bool norm = (F64 || MOM) && d.renorm && kind == kPassEvolve && !normed;
        // a resumed launch has gen0's masks on the device and walked gen0 + 1
        // in the prologue: its mask pass has no steps
        const bool none = kind == kPassPre && d.resume && !evaluate;
        const int pk = norm ? kPassNorm : none ? kPassNone : kind;
        const long pg = pk == kPassEvolve || pk == kPassNorm ? g : kind == kPassApply ? last : g0 - 1;
        const long mgen = evaluate || norm || none ? -1 : kind == kPassEvolve ? g + 1 : kind == kPassApply ? -1 : g0;
        const long wgen = evaluate || norm || none ? -1 : kind == kPassEvolve ? g + 2 : kind == kPassApply ? -1 : g0 + 1;
        run_pass<PREC, NT, NM, REM>(d, sh, lds, pk, pg, mgen, wgen);
        if (norm) {
            if (threadIdx.x < P) {
                double nrm = sh.norm[threadIdx.x];
                if constexpr (MOM) {
                    // the mixed kernel's workers left partial sums, in lane order
                    const double *q = lds + kOffPart + threadIdx.x * (64 / NM) * 2;
                    nrm = 0.0;
                    for (int j = 0; j < 64 / NM; j++) nrm += q[2 * j];
                }
                sh.rscale[g & 1][threadIdx.x] = d.zeroth / nrm;
            }
            __syncthreads();
            normed = true;
            continue;
        }
        normed = false;
        if (kind == kPassApply) break;

        if (kind == kPassInit || kind == kPassEvolve) {
            // Fitness (deac.cpp:1539-1555; initial fitness deac.cpp:957-976)
            // and greedy selection (deac.cpp:1556-1572).  Every thread forms
            // terms ((o_t - E_tm) / sigma_t)^2 over [t][m] in place of E (the
            // fused last term keeps its quotient), then one lane per member
            // adds them in timeslice order.  Mixed K.S: E = epar + dE 2^ex[t],
            // kept in registers until the selection decides which replace
            // their parent's.
            double ek[kEk];
            if (F64 && d.use_nfm) {
                // -1st-moment term from the model ISF, before the residuals
                // replace it (deac.cpp:961-965, 1543-1547): a fold of rounded
                // products over tau, the last fused when T is odd
                if (wave == kWorker0 && lane < P) {
                    const gcdouble *dt = (const gcdouble *)d.nfm_dt;
                    double mm = 0.0;
                    for (int t = 0; t < d.T; t++) {
                        const double e = lds[t * Ppad + lane];
                        mm = ((d.T & 1) && t == d.T - 1) ? fma(e, dt[t], mm) : mm + e * dt[t];
                    }
                    sh.nfm[lane] = mm;
                }
                __syncthreads();
            }
            if constexpr (MOM) {
                // the same term from the mixed model ISF E = epar + dE 2^ex[t]
                // (each value formed as the residuals below form it)
                if (d.use_nfm) {
                    if (wave == kWorker0 && lane < P) {
                        const gcdouble *dt = (const gcdouble *)d.nfm_dt;
                        double mm = 0.0;
                        for (int t = 0; t < d.T; t++) {
                            const double e = epar[t * Ppad + lane] + lds[t * Ppad + lane] * sh.kscale[t];
                            mm = ((d.T & 1) && t == d.T - 1) ? fma(e, dt[t], mm) : mm + e * dt[t];
                        }
                        sh.nfm[lane] = mm;
                    }
                    __syncthreads();
                }
            }
            if constexpr (F64) {
                for (int k = threadIdx.x; k < d.T * Ppad; k += kBatchThreads) {
                    const int t = k / Ppad;
                    if (k % Ppad < P) {
                        const double e = lds[k];
                        if (epar) epar[k] = e;
                        const double dv = (sh.obs[t] - e) / sh.sig[t];
                        lds[k] = t == tlast ? dv : dv * dv;
                    }
                }
            } else {
#pragma unroll
                for (int j = 0; j < kEk; j++) {
                    const int k = threadIdx.x + j * kBatchThreads;
                    ek[j] = 0.0;
                    if (k < d.T * Ppad && k % Ppad < P) {
                        const int t = k / Ppad;
                        const double e = epar[k] + lds[k] * sh.kscale[t];
                        ek[j] = e;
                        const double dv = (sh.obs[t] - e) / sh.sig[t];
                        lds[k] = t == tlast ? dv : dv * dv;
                    }
                }
            }
            __syncthreads();
            if (wave == kWorker0 && lane < Ppad) {
                int acc = 0;
                if (lane < P) {
                    double chi = 0.0;
                    for (int t = 0; t < d.T; t++) {
                        const double v = lds[t * Ppad + lane];
                        chi = t == tlast ? fma(v, v, chi) : chi + v;
                    }
                    double fit = chi / static_cast<double>(d.T);
                    if constexpr (MOM) {
                        // the trial's moments: the workers' partial sums, in lane order
                        if (kind == kPassEvolve && (d.use_first || d.use_third)) {
                            constexpr int H = 64 / NM;  // partial sums per member (C / 2)
                            const double *q = lds + kOffPart + lane * H * 2;
                            double a = 0.0, b = 0.0;
                            for (int j = 0; j < H; j++) {
                                a += q[2 * j];
                                b += q[2 * j + 1];
                            }
                            sh.mom[lane][0] = a;
                            sh.mom[lane][1] = b;
                        }
                    }
                    if ((F64 || MOM) && d.use_nfm) {  // deac.cpp:961-965, 1543-1547
                        const double q = (d.nfm_value - sh.nfm[lane]) / d.nfm_error;
                        fit = fma(q, q, fit);
                    }
                    if (d.use_first) {  // deac.cpp:969, 1551
                        const double dm = sh.mom[lane][0] - d.first_moment;
                        fit += (dm * dm) / d.first_moment;
                    }
                    if (d.use_third) {  // deac.cpp:972, 1554
                        const double q = (d.third_moment - sh.mom[lane][1]) / d.third_moment_error;
                        fit = fma(q, q, fit);
                    }
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
            if (evaluate) {
                // evaluation launch: the fitness of the stored population
                if (threadIdx.x < P) {
                    const int q = threadIdx.x / d.Pi;
                    ((gdouble *)d.fitness[q])[threadIdx.x - q * d.Pi] = sh.fitness[threadIdx.x];
                }
                return;
            }
            if constexpr (!F64) {
                // accepted trials' model ISF replaces their parent's
#pragma unroll
                for (int j = 0; j < kEk; j++) {
                    const int k = threadIdx.x + j * kBatchThreads;
                    if (k < d.T * Ppad && k % Ppad < P && sh.accepted[last % 3][k % Ppad]) epar[k] = ek[j];
                }
            }
        }

        // Top of generation g: best fitness, statistics, stop test
        // (deac.cpp:1145-1166); this generation's CR and F (deac.cpp:1180-1213).
        if (wave == kWorker0) {
            // lane q: instance q; the workgroup stops when any instance's
            // stop test fires (the others resume from generation g)
            const int Pi = d.Pi;
            bool stop = false;
            if (lane < d.Q && g < d.gen_end) {
                const int q = lane;
                const double *f = sh.fitness + q * Pi;
                double mn = f[0];
                for (int i = 0; i < Pi; i++)
                    if (f[i] < mn) mn = f[i];
                if (d.stats[q]) {
                    gdouble *st = (gdouble *)d.stats[q];
                    double sum = 0.0, sq = 0.0;
                    const int p4 = Pi & ~3;
                    for (int i = 0; i < Pi; i++) {
                        const double v = f[i];
                        sum += v;
                        sq = i < p4 ? sq + v * v : fma(v, v, sq);
                    }
                    const long gi = g - d.stats_base[q], len = d.stats_len[q];
                    st[gi] = sum / static_cast<double>(Pi);
                    st[len + gi] = mn;
                    st[2 * len + gi] = sq / static_cast<double>(Pi);
                }
                stop = mn <= d.stop[q];
                sh.qstop[q] = stop ? 1 : 0;
            }
            const bool any = __ballot(stop) != 0;
            if (lane == 0) sh.flag = g >= d.gen_end ? 2 : any ? 1 : 0;
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
        const int m = threadIdx.x, q = m / d.Pi, i = m - q * d.Pi;
        gdouble *fit = (gdouble *)d.fitness[q], *cr = (gdouble *)d.cr_old[q], *fw = (gdouble *)d.f_old[q];
        fit[i] = sh.fitness[m];
        if (last >= 0) ((__attribute__((address_space(1))) int *)d.accp[q])[i] = sh.accepted[last % 3][m];
        for (int sg = 0; sg < 2; sg++) {
            cr[i * 2 + sg] = sh.cr_old[m][sg];
            fw[i * 2 + sg] = sh.f_old[m][sg];
        }
    }
#ifdef DEAC_ROLE_TIMERS
    if (blockIdx.x == 0 && (threadIdx.x & 63) == 0) {
        const unsigned long long tend = __builtin_amdgcn_s_memtime();
        atomicAdd(&g_role[threadIdx.x >> 6][0], g_wait[threadIdx.x >> 6]);
        atomicAdd(&g_role[threadIdx.x >> 6][1], tend - tstart);
        atomicAdd(&g_role[threadIdx.x >> 6][2], static_cast<unsigned long long>(g - g0));
    }
#endif
    if (threadIdx.x < d.Q) {
        const int q = threadIdx.x;
        __attribute__((address_space(1))) uint64_t *rng = (__attribute__((address_space(1))) uint64_t *)d.rng[q];
        __attribute__((address_space(1))) long *status = (__attribute__((address_space(1))) long *)d.status[q];
        for (int w = 0; w < 4; w++) rng[w] = sh.gen_start[g & 3][q][w];
        status[0] = sh.flag == 1 && sh.qstop[q] ? 1 : 0;
        status[1] = g;
    }
}

}  // namespace deac
