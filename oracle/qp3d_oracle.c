/*
 * qp3d oracle: a CPU restatement of the reference's 3D pure-qP VTI forward-modeling
 * time loop (CPU_function.c:10-22,301-480, GPU_kernel.cu:733-812, modeling3D.c:49-125).
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline load this library, and only as the checker or as the timed CPU baseline.
 * The HIP library never links or calls it.
 *
 * Two instances are generated from qp3d_oracle_impl.h:
 *   *_f32  the reference's own arithmetic: float storage, float stencils, and double
 *          wherever the reference promotes (pow(), double literals). It is compiled with
 *          -ffp-contract=off so it is bit-identical to the reference's CPU kernels built
 *          in this container (oracle/_ref).
 *   *_f64  the same physics computed entirely in double (the FP64 oracle).
 *
 * Defects of the reference CPU path that this restatement resolves:
 *   - exchange() reads an uninitialized index (CPU_function.c:490,497): the GPU
 *     definition i = z*nx*ny + x*ny + y (GPU_kernel.cu:745) is used;
 *   - fields are never zeroed (modeling3D.c:34-48): every array is calloc'd;
 *   - the CPU T/B/F PML arrays are sized pml*ny*nz (modeling3D.c:38-40): the GPU sizes
 *     (3dPureqPModeling_single_shot.cu:144-186) are used.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define REAL float
#define SFX(n) n##_f32
#define SQRTR sqrtf
#include "qp3d_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRTR

#define REAL double
#define SFX(n) n##_f64
#define SQRTR sqrt
#include "qp3d_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRTR
