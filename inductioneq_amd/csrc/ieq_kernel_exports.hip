// Typed per-kernel entry points (include/inductioneq_kernels.h): each forwards its buffers,
// in the reference kernel's argument order, to ieq_run with the kernel's name, so names,
// checks and error codes are those of the list interface.
#include <cstdint>

#include "../../include/inductioneq_kernels.h"

namespace {

int run_named(ieq_program* prog, const char* name, const uint32_t g[3], const uint32_t l[3], void* const* args, int n,
              double* kernel_seconds) {
  const char* names[1] = {name};
  return ieq_run(prog, names, 1, g, l, args, n, kernel_seconds);
}

}  // namespace

#define IEQ_TYPED_HEAD(NAME) int ieq_##NAME(ieq_program* prog, const uint32_t g[3], const uint32_t l[3],
#define IEQ_TYPED1(NAME) \
  IEQ_TYPED_HEAD(NAME) void* a0, double* ks) { void* a[] = {a0}; return run_named(prog, #NAME, g, l, a, 1, ks); }
#define IEQ_TYPED2(NAME)                                                \
  IEQ_TYPED_HEAD(NAME) void* a0, void* a1, double* ks) {                \
    void* a[] = {a0, a1};                                               \
    return run_named(prog, #NAME, g, l, a, 2, ks);                      \
  }
#define IEQ_TYPED3(NAME)                                                \
  IEQ_TYPED_HEAD(NAME) void* a0, void* a1, void* a2, double* ks) {      \
    void* a[] = {a0, a1, a2};                                           \
    return run_named(prog, #NAME, g, l, a, 3, ks);                      \
  }
#define IEQ_TYPED6(NAME)                                                                       \
  IEQ_TYPED_HEAD(NAME) void* a0, void* a1, void* a2, void* a3, void* a4, void* a5, double* ks) { \
    void* a[] = {a0, a1, a2, a3, a4, a5};                                                      \
    return run_named(prog, #NAME, g, l, a, 6, ks);                                             \
  }
#define IEQ_TYPED7(NAME)                                                                                 \
  IEQ_TYPED_HEAD(NAME) void* a0, void* a1, void* a2, void* a3, void* a4, void* a5, void* a6, double* ks) { \
    void* a[] = {a0, a1, a2, a3, a4, a5, a6};                                                            \
    return run_named(prog, #NAME, g, l, a, 7, ks);                                                       \
  }

extern "C" {

IEQ_TYPED7(calc_time_3_args) IEQ_TYPED7(calc_u_3_args) IEQ_TYPED7(calc_curlB_rho_1_3args)
IEQ_TYPED7(calc_curlB_rho_2_3args) IEQ_TYPED7(calc_curlB_rho_3_3args) IEQ_TYPED7(SSPRK33_1)
IEQ_TYPED7(SSPRK33_2a) IEQ_TYPED7(SSPRK33_2b) IEQ_TYPED7(SSPRK33_3a) IEQ_TYPED7(SSPRK33_3b)
IEQ_TYPED7(SSPRK104_01) IEQ_TYPED7(SSPRK104_02a) IEQ_TYPED7(SSPRK104_02b) IEQ_TYPED7(SSPRK104_03a)
IEQ_TYPED7(SSPRK104_03b) IEQ_TYPED7(SSPRK104_04a) IEQ_TYPED7(SSPRK104_04b) IEQ_TYPED7(SSPRK104_05a)
IEQ_TYPED7(SSPRK104_05b) IEQ_TYPED7(SSPRK104_06) IEQ_TYPED7(SSPRK104_07) IEQ_TYPED7(SSPRK104_08a)
IEQ_TYPED7(SSPRK104_08b) IEQ_TYPED7(SSPRK104_09a) IEQ_TYPED7(SSPRK104_09b) IEQ_TYPED7(SSPRK104_10a)
IEQ_TYPED7(SSPRK104_10b) IEQ_TYPED7(SSPRK104_11a) IEQ_TYPED7(SSPRK104_11b)
IEQ_TYPED7(KennedyCarpenterLewis2R54C_1a) IEQ_TYPED7(KennedyCarpenterLewis2R54C_1b)
IEQ_TYPED7(KennedyCarpenterLewis2R54C_2a) IEQ_TYPED7(KennedyCarpenterLewis2R54C_2b)
IEQ_TYPED7(KennedyCarpenterLewis2R54C_3a) IEQ_TYPED7(KennedyCarpenterLewis2R54C_3b)
IEQ_TYPED7(KennedyCarpenterLewis2R54C_4a) IEQ_TYPED7(KennedyCarpenterLewis2R54C_4b)
IEQ_TYPED7(KennedyCarpenterLewis2R54C_5) IEQ_TYPED7(CalvoFrancoRandez2R64_1a)
IEQ_TYPED7(CalvoFrancoRandez2R64_1b) IEQ_TYPED7(CalvoFrancoRandez2R64_2a)
IEQ_TYPED7(CalvoFrancoRandez2R64_2b) IEQ_TYPED7(CalvoFrancoRandez2R64_3a)
IEQ_TYPED7(CalvoFrancoRandez2R64_3b) IEQ_TYPED7(CalvoFrancoRandez2R64_4a)
IEQ_TYPED7(CalvoFrancoRandez2R64_4b) IEQ_TYPED7(CalvoFrancoRandez2R64_5a)
IEQ_TYPED7(CalvoFrancoRandez2R64_5b) IEQ_TYPED7(CalvoFrancoRandez2R64_6) IEQ_TYPED6(calc_time_2_args)
IEQ_TYPED6(calc_u_2_args) IEQ_TYPED6(calc_curlB_rho_1_2args) IEQ_TYPED6(calc_curlB_rho_2_2args)
IEQ_TYPED6(CarpenterKennedy2N54_1a) IEQ_TYPED6(CarpenterKennedy2N54_1b) IEQ_TYPED6(CarpenterKennedy2N54_2a)
IEQ_TYPED6(CarpenterKennedy2N54_2b) IEQ_TYPED6(CarpenterKennedy2N54_3a) IEQ_TYPED6(CarpenterKennedy2N54_3b)
IEQ_TYPED6(CarpenterKennedy2N54_4a) IEQ_TYPED6(CarpenterKennedy2N54_4b) IEQ_TYPED6(CarpenterKennedy2N54_5a)
IEQ_TYPED6(CarpenterKennedy2N54_5b) IEQ_TYPED6(ToulorgeDesmet2N84F_1a) IEQ_TYPED6(ToulorgeDesmet2N84F_1b)
IEQ_TYPED6(ToulorgeDesmet2N84F_2a) IEQ_TYPED6(ToulorgeDesmet2N84F_2b) IEQ_TYPED6(ToulorgeDesmet2N84F_3a)
IEQ_TYPED6(ToulorgeDesmet2N84F_3b) IEQ_TYPED6(ToulorgeDesmet2N84F_4a) IEQ_TYPED6(ToulorgeDesmet2N84F_4b)
IEQ_TYPED6(ToulorgeDesmet2N84F_5a) IEQ_TYPED6(ToulorgeDesmet2N84F_5b) IEQ_TYPED6(ToulorgeDesmet2N84F_6a)
IEQ_TYPED6(ToulorgeDesmet2N84F_6b) IEQ_TYPED6(ToulorgeDesmet2N84F_7a) IEQ_TYPED6(ToulorgeDesmet2N84F_7b)
IEQ_TYPED6(ToulorgeDesmet2N84F_8a) IEQ_TYPED6(ToulorgeDesmet2N84F_8b) IEQ_TYPED3(calc_axpy)
IEQ_TYPED3(calc_xpay) IEQ_TYPED2(calc_div_divcleaning) IEQ_TYPED2(calc_laplace_divcleaning)
IEQ_TYPED2(projector_divcleaning) IEQ_TYPED3(dot_product) IEQ_TYPED1(init_b) IEQ_TYPED1(init_u)
IEQ_TYPED1(init_rho) IEQ_TYPED2(analytical_b) IEQ_TYPED2(norm2) IEQ_TYPED3(norm2_diff) IEQ_TYPED2(norm2_S)
IEQ_TYPED2(calc_div) IEQ_TYPED2(calc_curl) IEQ_TYPED2(calc_grad)

}  // extern "C"
