// Host buffer utilities of the 3D API (include/cal3DBuffer.h), semantics of
// OpenCAL-OMP/source/cal3DBuffer.c:19-250.  Host-side helpers for model
// set-up code; the stepping path never uses them.
#include <stdlib.h>
#include <string.h>

extern "C" {
#include "cal3DBuffer.h"
}

static inline size_t n3(int rows, int columns, int slices) { return (size_t)rows * columns * slices; }
static inline size_t at3(int rows, int columns, const CALCell3D& c)
{
	return ((size_t)c.k * rows + c.i) * columns + c.j;
}

#define BUFFER3_FUNCS(SUF, T)                                                                             \
	extern "C" T* calAllocBuffer3D##SUF(int rows, int columns, int slices)                                \
	{                                                                                                     \
		return (T*)malloc(sizeof(T) * n3(rows, columns, slices));                                         \
	}                                                                                                     \
	extern "C" void calDeleteBuffer3D##SUF(T* M) { free(M); }                                             \
	extern "C" void calCopyBuffer3D##SUF(T* src, T* dst, int rows, int columns, int slices)               \
	{                                                                                                     \
		memcpy(dst, src, sizeof(T) * n3(rows, columns, slices));                                          \
	}                                                                                                     \
	extern "C" void calCopyActiveCellsBuffer3D##SUF(T* src, T* dst, int rows, int columns, int slices,     \
	                                                struct CALCell3D* cells, int n)                       \
	{                                                                                                     \
		(void)slices;                                                                                     \
		for (int c = 0; c < n; ++c) {                                                                     \
			const size_t l = at3(rows, columns, cells[c]);                                                \
			if (dst[l] != src[l]) dst[l] = src[l];                                                        \
		}                                                                                                 \
	}                                                                                                     \
	extern "C" void calAddBuffer3D##SUF(T* a, T* b, T* d, int rows, int columns, int slices)              \
	{                                                                                                     \
		const size_t n = n3(rows, columns, slices);                                                       \
		for (size_t c = 0; c < n; ++c) d[c] = a[c] + b[c];                                                \
	}                                                                                                     \
	extern "C" void calSubtractBuffer3D##SUF(T* a, T* b, T* d, int rows, int columns, int slices)         \
	{                                                                                                     \
		const size_t n = n3(rows, columns, slices);                                                       \
		for (size_t c = 0; c < n; ++c) d[c] = a[c] - b[c];                                                \
	}                                                                                                     \
	extern "C" void calSetActiveCellsBuffer3D##SUF(T* M, int rows, int columns, int slices, T v,          \
	                                               struct CALCell3D* cells, int n)                        \
	{                                                                                                     \
		(void)slices;                                                                                     \
		for (int c = 0; c < n; ++c) M[at3(rows, columns, cells[c])] = v;                                  \
	}

BUFFER3_FUNCS(b, CALbyte)
BUFFER3_FUNCS(i, CALint)
BUFFER3_FUNCS(r, CALreal)

// cal3DBuffer.c:214-230: byte and int buffers are set with memset (so an int
// value other than 0/-1 fills every byte), reals element-wise.
extern "C" void calSetBuffer3Db(CALbyte* M, int rows, int columns, int slices, CALbyte v)
{
	memset(M, v, sizeof(CALbyte) * n3(rows, columns, slices));
}

extern "C" void calSetBuffer3Di(CALint* M, int rows, int columns, int slices, CALint v)
{
	memset(M, v, sizeof(CALint) * n3(rows, columns, slices));
}

extern "C" void calSetBuffer3Dr(CALreal* M, int rows, int columns, int slices, CALreal v)
{
	const size_t n = n3(rows, columns, slices);
	for (size_t c = 0; c < n; ++c) M[c] = v;
}
