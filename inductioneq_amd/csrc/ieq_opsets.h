/* Operator-set descriptors over the generated coefficient tables.
 *
 * An operator set is one SBP family: classical or boundary-optimised ("extended")
 * closures at order 2, 4 or 6.  Each row table has 2*nb+1 rows (nb left boundary
 * rows, the central row, nb right boundary rows) and w columns; column j weighs the
 * node at offset j-(w-1)/2 from the node being evaluated.  The -Laplacian tables
 * have their own nb and w; NS_D0 is not defined for the extended closures.
 * C and C++ compatible; shared by the C oracle and the HIP host library.
 */
#ifndef IEQ_OPSETS_H
#define IEQ_OPSETS_H

#include "ieq_tables.h"

typedef struct {
  const char *name;  /* "4thOrder", "6thOrderExtended", ... (reference module names) */
  int order;         /* formal interior order p = 2, 4, 6 */
  int nb;            /* boundary rows per side */
  int w;             /* first-derivative stencil width */
  int w_hod;         /* high-order dissipation row width */
  const double *diff;      /* (2nb+1) x w */
  const double *minv;      /* 2nb+1 */
  const double *dho;       /* (2nb+1) x w_hod */
  const double *diff_adj;  /* (2nb+1) x w, equals M^-1 D^T M */
  int lap_nb[3];           /* [0] WS_D0, [1] NS_D0, [2] WS_LN; 0 when absent */
  int lap_w[3];
  const double *lap[3];    /* NULL when absent */
} ieq_opset_desc;

#define IEQ_LAP_WS_D0 0
#define IEQ_LAP_NS_D0 1
#define IEQ_LAP_WS_LN 2

#define IEQ_OPSET_DESC(lo, UP, modname, NSD0)                                         \
  { modname, IEQ_##UP##_ORDER, IEQ_##UP##_NB, IEQ_##UP##_W, IEQ_##UP##_W_HOD,        \
    ieq_##lo##_diff, ieq_##lo##_minv, ieq_##lo##_dho, ieq_##lo##_diff_adj,           \
    { IEQ_##UP##_LAP_WS_D0_NB, IEQ_##UP##_LAP_NS_D0_NB, IEQ_##UP##_LAP_WS_LN_NB },   \
    { IEQ_##UP##_LAP_WS_D0_W, IEQ_##UP##_LAP_NS_D0_W, IEQ_##UP##_LAP_WS_LN_W },      \
    { ieq_##lo##_lap_ws_d0, NSD0, ieq_##lo##_lap_ws_ln } }

static const ieq_opset_desc ieq_opsets[IEQ_NUM_OPSETS] = {
  IEQ_OPSET_DESC(2nd, 2ND, "2ndOrder", ieq_2nd_lap_ns_d0),
  IEQ_OPSET_DESC(4th, 4TH, "4thOrder", ieq_4th_lap_ns_d0),
  IEQ_OPSET_DESC(6th, 6TH, "6thOrder", ieq_6th_lap_ns_d0),
  IEQ_OPSET_DESC(2nd_ext, 2ND_EXT, "2ndOrderExtended", 0),
  IEQ_OPSET_DESC(4th_ext, 4TH_EXT, "4thOrderExtended", 0),
  IEQ_OPSET_DESC(6th_ext, 6TH_EXT, "6thOrderExtended", 0),
};

#undef IEQ_OPSET_DESC

#endif /* IEQ_OPSETS_H */
