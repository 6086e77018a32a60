// valu_issue.hip -- issue-rate probe for the backprojector's roofline.
//
// Measures how many wave64 VALU instructions per ns one SIMD issues for
//   * single instruction kinds the backprojection loops use, and
//   * the per-voxel instruction mix of each loop (fp32, fp16 and bf16
//     staged texels), as the compiler emits it for backproject_lds_kernel,
// each in eight independent chains per lane, at 4, 5 and 8 waves per SIMD
// (the fp32 and 16-bit kernels run 4 and 5).  The mixes' rates are the
// ceilings the kernels' measured issue rates are compared with.  ALU
// instructions only (inline asm, no memory access); the result of every
// chain feeds a conditional store so that nothing is removed.
//
// Build: make -C 3dfilteredbackprojection-public_amd probes
// Run:   3dfilteredbackprojection-public_amd/probes/valu_issue
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>

namespace {

// About 1 ms per launch, so that the launch and the waves' ramp up and down
// are about 1% of the time.
constexpr int kIter = 16384;

#define CHECK(call)                                                                     \
    do {                                                                                \
        hipError_t e_ = (call);                                                         \
        if (e_ != hipSuccess) {                                                         \
            std::fprintf(stderr, "%s: %s\n", #call, hipGetErrorString(e_));             \
            std::exit(1);                                                               \
        }                                                                               \
    } while (0)

// A kernel running `body` kIter times on eight chains x0..x7 (operands %0..%7),
// with a float operand %8 and a packed-f16 / integer operand %9.
#define PROBE(NAME, BODY)                                                               \
    __global__ __launch_bounds__(256) void NAME(float *out, unsigned seed)              \
    {                                                                                   \
        float x0 = threadIdx.x, x1 = x0 + 1, x2 = x0 + 2, x3 = x0 + 3, x4 = x0 + 4,      \
              x5 = x0 + 5, x6 = x0 + 6, x7 = x0 + 7;                                      \
        const float a = __uint_as_float(seed | 0x3f000000u);                            \
        const unsigned h = seed * 2654435761u;                                          \
        for (int it = 0; it < kIter; ++it)                                              \
            asm volatile(BODY : "+v"(x0), "+v"(x1), "+v"(x2), "+v"(x3), "+v"(x4),        \
                         "+v"(x5), "+v"(x6), "+v"(x7) : "v"(a), "v"(h));                \
        const float s = x0 + x1 + x2 + x3 + x4 + x5 + x6 + x7;                           \
        if (s == 1234.5f) out[threadIdx.x] = s;                                         \
    }

#define EACH(OP) OP(0) OP(1) OP(2) OP(3) OP(4) OP(5) OP(6) OP(7)
#define FMAC(i) "v_fmac_f32 %" #i ", %8, %8\n"
#define SUBF(i) "v_sub_f32 %" #i ", %" #i ", %8\n"
#define MIX(i) "v_fma_mix_f32 %" #i ", %8, %9, %9 op_sel:[0,1,0] op_sel_hi:[0,1,1]\n"
#define SHL(i) "v_lshlrev_b32 %" #i ", 16, %9\n"
#define FRACT(i) "v_fract_f32 %" #i ", %" #i "\n"
#define CVTU(i) "v_cvt_u32_f32 %" #i ", %8\n"
#define MAD24(i) "v_mad_u32_u24 %" #i ", %9, %9, %" #i "\n"
#define FMAF(i) "v_fma_f32 %" #i ", %8, %8, %" #i "\n"
#define ADDF(i) "v_add_f32 %" #i ", %" #i ", %8\n"
#define ADDU(i) "v_add_u32 %" #i ", %9, %" #i "\n"
#define MAD16(i) "v_mad_u32_u16 %" #i ", %" #i ", %9, %" #i " op_sel:[1,0,0,0]\n"

PROBE(p_fmac, EACH(FMAC))
PROBE(p_fma, EACH(FMAF))
PROBE(p_mix, EACH(MIX))
PROBE(p_fract, EACH(FRACT))
PROBE(p_cvtu, EACH(CVTU))
PROBE(p_mad24, EACH(MAD24))
PROBE(p_shl, EACH(SHL))
PROBE(p_subf, EACH(SUBF))
PROBE(p_addf, EACH(ADDF))
PROBE(p_addu, EACH(ADDU))
PROBE(p_mad16, EACH(MAD16))

// Address part of one voxel of the fp32 and bf16 loops: row coordinate FMA,
// fract (row weight), cvt (row), mad24 (LDS address).
#define ADDR32 "v_fma_f32 %0, %8, %8, %0\n v_fract_f32 %1, %0\n v_cvt_u32_f32 %2, %0\n" \
               " v_mad_u32_u24 %3, %9, %9, %2\n"
// The same for the fp16 loop, whose rows are biased by 128: row
// coordinate FMA, fract, and the LDS address from the row's high half
// (v_mad_u32_u16 with op_sel).
#define ADDR16 "v_fma_f32 %0, %8, %8, %0\n v_fract_f32 %1, %0\n" \
               " v_mad_u32_u16 %3, %0, %9, %3 op_sel:[1,0,0,0]\n"
// fp16 value-and-step texels: two v_fma_mix_f32 lerps, two weighted FMAs
PROBE(p_loop_fp16, ADDR16 MIX(4) MIX(5) FMAC(6) FMAC(7))
// bf16 value-and-step texels: two shifts widen the values, two lerps, two FMAs
PROBE(p_loop_bf16, ADDR32 SHL(4) SHL(5) FMAC(6) FMAC(7) FMAC(4) FMAC(5))
// fp32 texels: two differences, two lerps, two weighted FMAs
PROBE(p_loop_fp32, ADDR32 SUBF(4) SUBF(5) FMAC(6) FMAC(7) FMAC(4) FMAC(5))

struct Probe {
    const char *name;
    void (*kernel)(float *, unsigned);
    int instructions;  // VALU instructions per chain step
};

}  // namespace

int main()
{
    float *out = nullptr;
    CHECK(hipMalloc(&out, 4096));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const Probe probes[] = {
        {"v_fmac_f32", p_fmac, 8},         {"v_fma_f32", p_fma, 8},
        {"v_fma_mix_f32", p_mix, 8},       {"v_fract_f32", p_fract, 8},
        {"v_cvt_u32_f32", p_cvtu, 8},      {"v_mad_u32_u24", p_mad24, 8},
        {"v_lshlrev_b32", p_shl, 8},       {"v_sub_f32", p_subf, 8},
        {"v_add_f32", p_addf, 8},          {"v_add_u32", p_addu, 8},
        {"v_mad_u32_u16", p_mad16, 8},     {"loop mix fp16", p_loop_fp16, 7},
        {"loop mix bf16", p_loop_bf16, 10}, {"loop mix fp32", p_loop_fp32, 10}};
    std::printf("VALU issue rate, wave64 instructions per ns per SIMD (%d CUs, 4 SIMDs each)\n",
                cus);
    std::printf("%-16s %10s %10s %10s\n", "", "4 waves", "5 waves", "8 waves");
    for (const Probe &p : probes) {
        std::printf("%-16s", p.name);
        for (int waves : {4, 5, 8}) {
            const int blocks = cus * waves;  // 4-wave workgroups: `waves` per SIMD
            float best = 1e30f;
            for (int rep = 0; rep < 3; ++rep) {
                CHECK(hipEventRecord(e0));
                hipLaunchKernelGGL(p.kernel, dim3(blocks), dim3(256), 0, 0, out, 3u);
                CHECK(hipGetLastError());
                CHECK(hipEventRecord(e1));
                CHECK(hipEventSynchronize(e1));
                float ms = 0.0f;
                CHECK(hipEventElapsedTime(&ms, e0, e1));
                if (rep > 0 && ms < best) best = ms;
            }
            const double instr = (double)blocks * 4 * kIter * p.instructions;
            std::printf(" %10.3f", instr / (4.0 * cus) / (best * 1e6));
        }
        std::printf("\n");
    }
    CHECK(hipFree(out));
    return 0;
}
