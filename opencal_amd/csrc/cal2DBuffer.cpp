// Host buffer utilities (include/cal2DBuffer.h), semantics of
// OpenCAL-OMP/source/cal2DBuffer.c:20-259.  Host-side helpers for model set-up
// code; the stepping path never uses them.
#include <stdlib.h>
#include <string.h>

extern "C" {
#include "cal2DBuffer.h"
}

template <typename T>
static T* alloc2(int rows, int columns) { return (T*)malloc(sizeof(T) * (size_t)rows * (size_t)columns); }

template <typename T>
static void copy_active(const T* src, T* dst, int columns, const CALCell2D* cells, int n)
{
	for (int k = 0; k < n; ++k) {
		size_t c = (size_t)cells[k].i * columns + cells[k].j;
		if (dst[c] != src[c]) dst[c] = src[c];
	}
}

template <typename T>
static void set_active(T* M, int columns, T v, const CALCell2D* cells, int n)
{
	for (int k = 0; k < n; ++k) M[(size_t)cells[k].i * columns + cells[k].j] = v;
}

template <typename T>
static void add2(const T* a, const T* b, T* d, int rows, int columns)
{
	size_t n = (size_t)rows * columns;
	for (size_t c = 0; c < n; ++c) d[c] = a[c] + b[c];
}

template <typename T>
static void sub2(const T* a, const T* b, T* d, int rows, int columns)
{
	size_t n = (size_t)rows * columns;
	for (size_t c = 0; c < n; ++c) d[c] = a[c] - b[c];
}

#define BUFFER_FUNCS(SUF, T)                                                                              \
	extern "C" T* calAllocBuffer2D##SUF(int rows, int columns) { return alloc2<T>(rows, columns); }       \
	extern "C" void calDeleteBuffer2D##SUF(T* M) { free(M); }                                             \
	extern "C" void calCopyBuffer2D##SUF(T* src, T* dst, int rows, int columns)                           \
	{                                                                                                     \
		memcpy(dst, src, sizeof(T) * (size_t)rows * columns);                                             \
	}                                                                                                     \
	extern "C" void calCopyActiveCellsBuffer2D##SUF(T* src, T* dst, int rows, int columns,                 \
	                                                struct CALCell2D* cells, int n)                       \
	{                                                                                                     \
		(void)rows;                                                                                       \
		copy_active<T>(src, dst, columns, cells, n);                                                      \
	}                                                                                                     \
	extern "C" void calAddBuffer2D##SUF(T* a, T* b, T* d, int rows, int columns) { add2<T>(a, b, d, rows, columns); } \
	extern "C" void calSubtractBuffer2D##SUF(T* a, T* b, T* d, int rows, int columns) { sub2<T>(a, b, d, rows, columns); } \
	extern "C" void calSetActiveCellsBuffer2D##SUF(T* M, int rows, int columns, T v, struct CALCell2D* cells, int n) \
	{                                                                                                     \
		(void)rows;                                                                                       \
		set_active<T>(M, columns, v, cells, n);                                                           \
	}

BUFFER_FUNCS(b, CALbyte)
BUFFER_FUNCS(i, CALint)
BUFFER_FUNCS(r, CALreal)

// cal2DBuffer.c:223-238: byte and int buffers are set with memset (so an int
// value other than 0/-1 fills every byte), reals element-wise.
extern "C" void calSetBuffer2Db(CALbyte* M, int rows, int columns, CALbyte v)
{
	memset(M, v, sizeof(CALbyte) * (size_t)rows * columns);
}

extern "C" void calSetBuffer2Di(CALint* M, int rows, int columns, CALint v)
{
	memset(M, v, sizeof(CALint) * (size_t)rows * columns);
}

extern "C" void calSetBuffer2Dr(CALreal* M, int rows, int columns, CALreal v)
{
	size_t n = (size_t)rows * columns;
	for (size_t c = 0; c < n; ++c) M[c] = v;
}
