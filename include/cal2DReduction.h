/* Global reductions of a 2D substate (OpenCAL-OMP/include/cal2DReduction.h,
 * OpenCAL-OMP/source/cal2DReduction.c:18-641), computed on the GPU.
 *
 * Every function folds op over the `current` layer of all rows*columns
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
#ifndef cal2DReduction_h
#define cal2DReduction_h

#include <cal2D.h>

#ifdef __cplusplus
extern "C" {
#endif

CALbyte calReductionOperation2Db(struct CALModel2D* model, struct CALSubstate2Db* substate, enum REDUCTION_OPERATION operation);
CALint calReductionOperation2Di(struct CALModel2D* model, struct CALSubstate2Di* substate, enum REDUCTION_OPERATION operation);
CALreal calReductionOperation2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate, enum REDUCTION_OPERATION operation);

CALbyte calReductionComputeMax2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeMax2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeMax2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeMin2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeMin2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeMin2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeSum2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeSum2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeSum2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
/* As in the reference (cal2DReduction.c:563-571): the byte variant is the
 * product of the cells, the int and real variants return their SUM. */
CALbyte calReductionComputeProd2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeProd2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeProd2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeLogicalAnd2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeLogicalAnd2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeLogicalAnd2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeBinaryAnd2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeBinaryAnd2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeBinaryAnd2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeLogicalOr2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeLogicalOr2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeLogicalOr2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeBinaryOr2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeBinaryOr2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeBinaryOr2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeLogicalXor2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeLogicalXor2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeLogicalXor2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);
CALbyte calReductionComputeBinaryXor2Db(struct CALModel2D* model, struct CALSubstate2Db* substate);
CALint calReductionComputeBinaryXor2Di(struct CALModel2D* model, struct CALSubstate2Di* substate);
CALreal calReductionComputeBinaryXor2Dr(struct CALModel2D* model, struct CALSubstate2Dr* substate);

/* current[index] of the substate (cal2DReduction.h getValue2D*AtIndex). */
CALbyte getValue2DbAtIndex(struct CALSubstate2Db* substate, CALint index);
CALint getValue2DiAtIndex(struct CALSubstate2Di* substate, CALint index);
CALreal getValue2DrAtIndex(struct CALSubstate2Dr* substate, CALint index);

#ifdef __cplusplus
}
#endif
#endif
