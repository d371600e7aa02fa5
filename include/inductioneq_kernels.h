/* Typed entry points of libinductioneq_hip.so: one function per reference OpenCL kernel.
 *
 * ieq_<name>(prog, g, l, <the kernel's buffers in the reference order>, kernel_seconds) runs
 * the kernel <name> exactly as ieq_run(prog, &"<name>", 1, g, l, args, n, kernel_seconds)
 * does (same checks, same error codes, same device layout: vector fields are the SoA
 * fields of inductioneq.h, scalars such as the CG coefficient are 1-element buffers).  They
 * are the per-kernel calls a plain C host writes instead of building argument arrays, with
 * the parameter names of the reference kernel signatures (file:line above each one).
 * `l` matters only for the reductions (norm2*, dot_product: one output per l[0] nodes) and
 * may be NULL elsewhere; kernel_seconds may be NULL.
 */
#ifndef INDUCTIONEQ_KERNELS_H
#define INDUCTIONEQ_KERNELS_H

#include "inductioneq.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Runge-Kutta stages, time and velocity updates and J = curl B / rho of the three-register integrators */
/* kernel_time_integrator.cl:65 */
int ieq_calc_time_3_args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                         void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                         void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:112 */
int ieq_calc_u_3_args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                      void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                      void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:168 */
int ieq_calc_curlB_rho_1_3args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_b3,
                               void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                               double *kernel_seconds);
/* kernel_time_integrator.cl:182 */
int ieq_calc_curlB_rho_2_3args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_b3,
                               void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                               double *kernel_seconds);
/* kernel_time_integrator.cl:196 */
int ieq_calc_curlB_rho_3_3args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_b3,
                               void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                               double *kernel_seconds);
/* kernel_time_integrator.cl:224 */
int ieq_SSPRK33_1(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                  void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                  void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:239 */
int ieq_SSPRK33_2a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                   void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                   void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:256 */
int ieq_SSPRK33_2b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                   void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                   void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:272 */
int ieq_SSPRK33_3a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                   void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                   void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:289 */
int ieq_SSPRK33_3b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                   void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                   void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:315 */
int ieq_SSPRK104_01(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                    void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                    void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:330 */
int ieq_SSPRK104_02a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:347 */
int ieq_SSPRK104_02b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:362 */
int ieq_SSPRK104_03a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:379 */
int ieq_SSPRK104_03b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:394 */
int ieq_SSPRK104_04a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:411 */
int ieq_SSPRK104_04b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:426 */
int ieq_SSPRK104_05a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:443 */
int ieq_SSPRK104_05b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:458 */
int ieq_SSPRK104_06(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                    void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                    void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:479 */
int ieq_SSPRK104_07(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                    void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                    void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:494 */
int ieq_SSPRK104_08a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:511 */
int ieq_SSPRK104_08b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:526 */
int ieq_SSPRK104_09a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:543 */
int ieq_SSPRK104_09b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:558 */
int ieq_SSPRK104_10a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:575 */
int ieq_SSPRK104_10b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:590 */
int ieq_SSPRK104_11a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:607 */
int ieq_SSPRK104_11b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b1, void *d_field_b2, void *d_field_b3, void *d_field_curlB_rho,
                     void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:630 */
int ieq_KennedyCarpenterLewis2R54C_1a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:643 */
int ieq_KennedyCarpenterLewis2R54C_1b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:668 */
int ieq_KennedyCarpenterLewis2R54C_2a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:681 */
int ieq_KennedyCarpenterLewis2R54C_2b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:706 */
int ieq_KennedyCarpenterLewis2R54C_3a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:719 */
int ieq_KennedyCarpenterLewis2R54C_3b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:738 */
int ieq_KennedyCarpenterLewis2R54C_4a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:751 */
int ieq_KennedyCarpenterLewis2R54C_4b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                      void *d_field_b1, void *d_field_b2, void *d_field_db,
                                      void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                      void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:776 */
int ieq_KennedyCarpenterLewis2R54C_5(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                     void *d_field_b1, void *d_field_b2, void *d_field_db,
                                     void *d_field_curlB_rho, void *d_field_u, void *d_field_rho,
                                     void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:801 */
int ieq_CalvoFrancoRandez2R64_1a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:814 */
int ieq_CalvoFrancoRandez2R64_1b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:839 */
int ieq_CalvoFrancoRandez2R64_2a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:852 */
int ieq_CalvoFrancoRandez2R64_2b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:877 */
int ieq_CalvoFrancoRandez2R64_3a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:890 */
int ieq_CalvoFrancoRandez2R64_3b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:915 */
int ieq_CalvoFrancoRandez2R64_4a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:928 */
int ieq_CalvoFrancoRandez2R64_4b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:953 */
int ieq_CalvoFrancoRandez2R64_5a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:966 */
int ieq_CalvoFrancoRandez2R64_5b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field_b1, void *d_field_b2, void *d_field_db,
                                 void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                 double *kernel_seconds);
/* kernel_time_integrator.cl:991 */
int ieq_CalvoFrancoRandez2R64_6(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_db,
                                void *d_field_curlB_rho, void *d_field_u, void *d_field_rho, void *time,
                                double *kernel_seconds);

/* Runge-Kutta stages, time and velocity updates and J = curl B / rho of the two-register (2N) integrators */
/* kernel_time_integrator.cl:42 */
int ieq_calc_time_2_args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                         void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho, void *d_field_u,
                         void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:93 */
int ieq_calc_u_2_args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                      void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho, void *d_field_u,
                      void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:136 */
int ieq_calc_curlB_rho_1_2args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:150 */
int ieq_calc_curlB_rho_2_2args(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1023 */
int ieq_CarpenterKennedy2N54_1a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1036 */
int ieq_CarpenterKennedy2N54_1b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1059 */
int ieq_CarpenterKennedy2N54_2a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1075 */
int ieq_CarpenterKennedy2N54_2b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1098 */
int ieq_CarpenterKennedy2N54_3a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1114 */
int ieq_CarpenterKennedy2N54_3b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1137 */
int ieq_CarpenterKennedy2N54_4a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1153 */
int ieq_CarpenterKennedy2N54_4b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1176 */
int ieq_CarpenterKennedy2N54_5a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1192 */
int ieq_CarpenterKennedy2N54_5b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                                void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1216 */
int ieq_ToulorgeDesmet2N84F_1a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1229 */
int ieq_ToulorgeDesmet2N84F_1b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1252 */
int ieq_ToulorgeDesmet2N84F_2a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1268 */
int ieq_ToulorgeDesmet2N84F_2b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1291 */
int ieq_ToulorgeDesmet2N84F_3a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1307 */
int ieq_ToulorgeDesmet2N84F_3b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1330 */
int ieq_ToulorgeDesmet2N84F_4a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1346 */
int ieq_ToulorgeDesmet2N84F_4b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1369 */
int ieq_ToulorgeDesmet2N84F_5a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1385 */
int ieq_ToulorgeDesmet2N84F_5b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1408 */
int ieq_ToulorgeDesmet2N84F_6a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1424 */
int ieq_ToulorgeDesmet2N84F_6b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1447 */
int ieq_ToulorgeDesmet2N84F_7a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1463 */
int ieq_ToulorgeDesmet2N84F_7b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1486 */
int ieq_ToulorgeDesmet2N84F_8a(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);
/* kernel_time_integrator.cl:1502 */
int ieq_ToulorgeDesmet2N84F_8b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                               void *d_field_b1, void *d_field_b2, void *d_field_curlB_rho,
                               void *d_field_u, void *d_field_rho, void *time, double *kernel_seconds);

/* Initial and analytic fields, diagnostics and operators, divergence cleaning and CG helpers */
/* kernel_CG.cl:8 */
int ieq_calc_axpy(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                  void *d_x, void *d_y, void *a, double *kernel_seconds);
/* kernel_CG.cl:21 */
int ieq_calc_xpay(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                  void *d_x, void *d_y, void *a, double *kernel_seconds);
/* kernel_div_cleaning.cl:60 */
int ieq_calc_div_divcleaning(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                             void *d_field, void *d_field_div, double *kernel_seconds);
/* kernel_div_cleaning.cl:75 */
int ieq_calc_laplace_divcleaning(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                                 void *d_field, void *d_field_laplace, double *kernel_seconds);
/* kernel_div_cleaning.cl:96 */
int ieq_projector_divcleaning(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                              void *d_field_b, void *d_field_phi, double *kernel_seconds);
/* kernel_dot_product.cl:47 */
int ieq_dot_product(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                    void *a_vec, void *b_vec, void *output, double *kernel_seconds);
/* kernel_init.cl:3 */
int ieq_init_b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
               void *d_field_b, double *kernel_seconds);
/* kernel_init.cl:15 */
int ieq_init_u(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
               void *d_field_u, double *kernel_seconds);
/* kernel_init.cl:27 */
int ieq_init_rho(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                 void *d_field_rho, double *kernel_seconds);
/* kernel_init.cl:39 */
int ieq_analytical_b(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                     void *d_field_b, void *time, double *kernel_seconds);
/* kernel_operator.cl:11 */
int ieq_norm2(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
              void *d_field, void *output, double *kernel_seconds);
/* kernel_operator.cl:55 */
int ieq_norm2_diff(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                   void *d_field_1, void *d_field_2, void *output, double *kernel_seconds);
/* kernel_operator.cl:99 */
int ieq_norm2_S(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                void *d_field, void *output, double *kernel_seconds);
/* kernel_operator.cl:145 */
int ieq_calc_div(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                 void *d_field, void *d_field_div, double *kernel_seconds);
/* kernel_operator.cl:158 */
int ieq_calc_curl(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                  void *d_field, void *d_field_curl, double *kernel_seconds);
/* kernel_operator.cl:170 */
int ieq_calc_grad(ieq_program *prog, const uint32_t g[3], const uint32_t l[3],
                  void *d_field, void *d_field_grad, double *kernel_seconds);

#ifdef __cplusplus
}
#endif

#endif /* INDUCTIONEQ_KERNELS_H */
