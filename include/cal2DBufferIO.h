/* Matrix text IO on plain row-major buffers (OpenCAL-OMP/include/
 * cal2DBufferIO.h:22-72): whitespace-separated values, one matrix row per
 * text line; reals are written with "%.9f" as the OpenMP reference does
 * (OpenCAL-OMP/source/cal2DBufferIO.c:142).  These functions touch only the
 * buffers they are given (host memory). */
#ifndef cal2DBufferIO_h
#define cal2DBufferIO_h

#include <stdio.h>
#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

void calfLoadMatrix2Db(CALbyte* M, int rows, int columns, FILE* f);
void calfLoadMatrix2Di(CALint* M, int rows, int columns, FILE* f);
void calfLoadMatrix2Dr(CALreal* M, int rows, int columns, FILE* f);
CALbyte calLoadMatrix2Db(CALbyte* M, int rows, int columns, char* path);
CALbyte calLoadMatrix2Di(CALint* M, int rows, int columns, char* path);
CALbyte calLoadMatrix2Dr(CALreal* M, int rows, int columns, char* path);
void calfSaveMatrix2Db(CALbyte* M, int rows, int columns, FILE* f);
void calfSaveMatrix2Di(CALint* M, int rows, int columns, FILE* f);
void calfSaveMatrix2Dr(CALreal* M, int rows, int columns, FILE* f);
CALbyte calSaveMatrix2Db(CALbyte* M, int rows, int columns, char* path);
CALbyte calSaveMatrix2Di(CALint* M, int rows, int columns, char* path);
CALbyte calSaveMatrix2Dr(CALreal* M, int rows, int columns, char* path);

#ifdef __cplusplus
}
#endif
#endif
