/* Global reductions of a 3D substate (OpenCAL-OMP/include/cal3DReduction.h,
 * OpenCAL-OMP/source/cal3DReduction.c:18-642), computed on the GPU.
 *
 * Every function folds op over the `current` layer of all rows*columns*slices
 * cells in index order, starting from cell 0, as the reference's serial
 * build would if its loop covered the grid: the reference ends its last
 * chunk at `rows` instead of the cell count and, with OpenMP, starts every
 * thread's partial from cell 0, so its own result depends on the thread
 * count.  On a single cell the result is that cell's value.
 *  - integer and byte reductions, MAX and MIN (ties and NaN as the serial
 *    fold decides them) and the logical/bitwise ones are exact;
 *  - SUM of reals is compensated (Neumaier) in a fixed tree: deterministic,
 *    within about one rounding of the exact sum; PROD of reals is a fixed
 *    tree of products (deterministic, not the serial rounding);
 *  - BINARY_* of reals convert each value to CALint first as the reference
 *    does; values outside the int range (undefined in C) convert to INT_MIN
 *    as on the x86 build.
 * The substate is read where it lives: on the GPU during a device run (no
 * transfer), otherwise its host array is copied to the GPU for the call. */
#ifndef cal3DReduction_h
#define cal3DReduction_h

#include <cal3D.h>

#ifdef __cplusplus
extern "C" {
#endif

CALbyte calReductionOperation3Db(struct CALModel3D* model, struct CALSubstate3Db* substate, enum REDUCTION_OPERATION operation);
CALint calReductionOperation3Di(struct CALModel3D* model, struct CALSubstate3Di* substate, enum REDUCTION_OPERATION operation);
CALreal calReductionOperation3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate, enum REDUCTION_OPERATION operation);

CALbyte calReductionComputeMax3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeMax3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeMax3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeMin3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeMin3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeMin3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeSum3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeSum3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeSum3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
/* As in the reference (cal3DReduction.c:563-571): the byte variant is the
 * product of the cells, the int and real variants return their SUM. */
CALbyte calReductionComputeProd3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeProd3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeProd3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeLogicalAnd3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeLogicalAnd3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeLogicalAnd3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeBinaryAnd3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeBinaryAnd3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeBinaryAnd3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeLogicalOr3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeLogicalOr3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeLogicalOr3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeBinaryOr3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeBinaryOr3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeBinaryOr3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeLogicalXor3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeLogicalXor3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeLogicalXor3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);
CALbyte calReductionComputeBinaryXor3Db(struct CALModel3D* model, struct CALSubstate3Db* substate);
CALint calReductionComputeBinaryXor3Di(struct CALModel3D* model, struct CALSubstate3Di* substate);
CALreal calReductionComputeBinaryXor3Dr(struct CALModel3D* model, struct CALSubstate3Dr* substate);

/* current[index] of the substate (cal3DReduction.h getValue3D*AtIndex). */
CALbyte getValue3DbAtIndex(struct CALSubstate3Db* substate, CALint index);
CALint getValue3DiAtIndex(struct CALSubstate3Di* substate, CALint index);
CALreal getValue3DrAtIndex(struct CALSubstate3Dr* substate, CALint index);

#ifdef __cplusplus
}
#endif
#endif
