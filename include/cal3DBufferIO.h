/* Text IO of 3D host buffers (OpenCAL-OMP/include/cal3DBufferIO.h:22-72):
 * whitespace-separated values, slice by slice, one row per text line and a
 * blank line after each slice; reals are written with "%f" as
 * cal3DBufferIO.c:144-160 does.  These functions touch only the buffers they
 * are given (host memory). */
#ifndef cal3DBufferIO_h
#define cal3DBufferIO_h

#include <stdio.h>
#include <calCommon.h>

#ifdef __cplusplus
extern "C" {
#endif

void calfLoadBuffer3Db(CALbyte* M, int rows, int columns, int slices, FILE* f);
void calfLoadBuffer3Di(CALint* M, int rows, int columns, int slices, FILE* f);
void calfLoadBuffer3Dr(CALreal* M, int rows, int columns, int slices, FILE* f);
CALbyte calLoadBuffer3Db(CALbyte* M, int rows, int columns, int slices, char* path);
CALbyte calLoadBuffer3Di(CALint* M, int rows, int columns, int slices, char* path);
CALbyte calLoadBuffer3Dr(CALreal* M, int rows, int columns, int slices, char* path);
void calfSaveBuffer3Db(CALbyte* M, int rows, int columns, int slices, FILE* f);
void calfSaveBuffer3Di(CALint* M, int rows, int columns, int slices, FILE* f);
void calfSaveBuffer3Dr(CALreal* M, int rows, int columns, int slices, FILE* f);
CALbyte calSaveBuffer3Db(CALbyte* M, int rows, int columns, int slices, char* path);
CALbyte calSaveBuffer3Di(CALint* M, int rows, int columns, int slices, char* path);
CALbyte calSaveBuffer3Dr(CALreal* M, int rows, int columns, int slices, char* path);

#ifdef __cplusplus
}
#endif
#endif
