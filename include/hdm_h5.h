/*
 * hdm_h5.h - HDF5 volume loader of libhdm.so (C ABI).
 *
 * Replaces the reference's benchmark loader loadHDFIntoBoolArr
 * (forBench/hdf5Manag.cu:44-83), which reads a dataset of an HDF5 file into a
 * bool array through the HDF5 C++ library, and its driver loadHDF
 * (forBench/hdf5Manag.cu:196-220), which loads two such datasets of one file
 * (512 x 512 x 826 flattened booleans) and computes their Hausdorff distance.
 * The reader is the library's own (no libhdf5): it follows the HDF5 file
 * format specification for what HDF5 1.8-1.14 and h5py write for plain
 * numeric datasets:
 *   superblock versions 0-3; object header versions 1 and 2; groups as
 *   symbol tables (B-tree + local heap) or compact link messages; dataset
 *   layouts compact, contiguous and chunked (B-tree v1, single-chunk,
 *   implicit and fixed-array chunk indexes); filters deflate, shuffle and
 *   fletcher32; element types integer (1/2/4/8 bytes, either byte order),
 *   enumeration over an integer (h5py's bool), bitfield and float (4/8).
 * Anything else returns HDM_E_UNSUPPORTED rather than guessing.
 */
#ifndef HDM_H5_H
#define HDM_H5_H

#include <stdint.h>

#include "hdm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define HDM_E_IO (-10)          /* the file cannot be opened or read            */
#define HDM_E_UNSUPPORTED (-11) /* an HDF5 feature the reader does not handle   */
#define HDM_E_FORMAT (-12)      /* not an HDF5 file, or a corrupt structure     */
#define HDM_E_NOTFOUND (-13)    /* no dataset at that path                      */

/* element classes */
#define HDM_H5_INT 0   /* signed integer                          */
#define HDM_H5_UINT 1  /* unsigned integer or bitfield            */
#define HDM_H5_FLOAT 2 /* IEEE float                              */
#define HDM_H5_ENUM 3  /* enumeration over an integer (h5py bool) */

/* layouts */
#define HDM_H5_COMPACT 0
#define HDM_H5_CONTIGUOUS 1
#define HDM_H5_CHUNKED 2

typedef struct hdm_h5_info {
  int32_t rank;        /* 0 (scalar) .. 8                                  */
  int32_t type_class;  /* HDM_H5_*                                         */
  int32_t elem_size;   /* bytes per element                                */
  int32_t layout;      /* HDM_H5_COMPACT / CONTIGUOUS / CHUNKED            */
  uint64_t dims[8];    /* slowest first (C order), dims[rank..7] = 0       */
  uint64_t n_elems;    /* product of dims                                  */
  uint64_t chunk[8];   /* chunk dims when chunked                          */
  int32_t n_filters;   /* filters in the pipeline                          */
  int32_t big_endian;  /* element byte order                               */
} hdm_h5_info;

/* read conversions */
#define HDM_H5_AS_U8 0      /* uint8; every value must lie in 0..255         */
#define HDM_H5_AS_I32 1     /* int32; every value must fit                   */
#define HDM_H5_AS_NONZERO 2 /* uint8 0 / 1: 1 where the element is not zero,
                               the reference's read as NATIVE_HBOOL
                               (forBench/hdf5Manag.cu:76)                    */

/* Shape, type and layout of dataset `dataset` ("/a/b" or "b") in `path`. */
int hdm_h5_info_get(const char *path, const char *dataset, hdm_h5_info *out);

/* Reads the whole dataset into dst (host memory, n_elems elements of the
 * conversion's type), in the file's C order.  threads <= 0: the library
 * picks (chunk decoding runs in parallel).  HDM_E_INVALID when dst_elems
 * differs from the dataset's element count or a value does not fit. */
int hdm_h5_read(const char *path, const char *dataset, int conversion, void *dst,
                uint64_t dst_elems, int threads);

/* The reference's loadHDF (forBench/hdf5Manag.cu:196-220): two datasets of
 * nx * ny * nz elements (any rank; x fastest), on where non-zero, then
 * hdm_hausdorff_u8 with label 1 and the given options (opts->ptr_kind is
 * ignored: the volumes are loaded to the host and copied).  load_ms (may be
 * NULL) receives the milliseconds spent reading both datasets. */
int hdm_hausdorff_h5(const char *gold_path, const char *gold_dataset, const char *segm_path,
                     const char *segm_dataset, int nx, int ny, int nz, const hdm_opts *opts,
                     hdm_result *out, double *load_ms);

#ifdef __cplusplus
}
#endif
#endif /* HDM_H5_H */
