/*
 * make_h5.c - writes the HDF5 fixtures of the loader tests with the HDF5
 * library itself (libhdf5 1.10 as found in the build container), so that the
 * library's own reader (housdorffhdf5cmake_amd/csrc/hdm_h5.cpp) is checked
 * against files it did not write.  Run by make_h5.sh; the files it writes are
 * committed under tests/golden/h5/.
 *
 * Element values follow closed formulas of the linear index i (C order), which
 * the tests restate: u8 (37 i + 11) mod 256, i32 7919 i - 100000, u16 i mod
 * 300, f32 1.0 where i mod 3 == 0, bool from two spheres (sphere_value).
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define CHECK(x)                                                   \
  do {                                                             \
    if ((x) < 0) {                                                 \
      fprintf(stderr, "HDF5 call failed at line %d\n", __LINE__); \
      exit(1);                                                     \
    }                                                              \
  } while (0)

static unsigned char f_u8(size_t i) { return (unsigned char)((37 * i + 11) % 256); }
static int f_i32(size_t i) { return (int)(7919 * (long)i - 100000); }
static unsigned short f_u16(size_t i) { return (unsigned short)(i % 300); }
static float f_f32(size_t i) { return i % 3 == 0 ? 1.0f : 0.0f; }

/* the bool pair: x fastest over (nz, ny, nx) = (20, 30, 40); gold a sphere of
 * radius 7 at (20, 15, 10), segm one of radius 7 at (22, 16, 10) plus the
 * voxel (3, 4, 5) */
static int sphere_value(int which, size_t i) {
  const int nx = 40, ny = 30;
  const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / (nx * ny));
  const int cx = which ? 22 : 20, cy = which ? 16 : 15, cz = 10;
  const int d2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
  if (which && x == 3 && y == 4 && z == 5) return 1;
  return d2 <= 49;
}

enum { U8, I32, U16, F32, I32BE, BOOL_GOLD, BOOL_SEGM, B8 };

static void *values(int kind, size_t n) {
  size_t es = kind == U8 || kind == B8 ? 1 : kind == U16 ? 2 : (kind == BOOL_GOLD || kind == BOOL_SEGM) ? 1 : 4;
  void *b = malloc(n * es);
  for (size_t i = 0; i < n; ++i) {
    switch (kind) {
      case U8:
      case B8: ((unsigned char *)b)[i] = f_u8(i); break;
      case I32:
      case I32BE: ((int *)b)[i] = f_i32(i); break;
      case U16: ((unsigned short *)b)[i] = f_u16(i); break;
      case F32: ((float *)b)[i] = f_f32(i); break;
      case BOOL_GOLD: ((signed char *)b)[i] = (signed char)sphere_value(0, i); break;
      case BOOL_SEGM: ((signed char *)b)[i] = (signed char)sphere_value(1, i); break;
    }
  }
  return b;
}

static hid_t bool_type(void) {
  /* what h5py writes for numpy bool: an enum over int8 */
  hid_t t = H5Tenum_create(H5T_NATIVE_INT8);
  signed char v0 = 0, v1 = 1;
  CHECK(H5Tenum_insert(t, "FALSE", &v0));
  CHECK(H5Tenum_insert(t, "TRUE", &v1));
  return t;
}

/* creates dataset `name` of `kind` with the given shape and creation list */
static void put(hid_t loc, const char *name, int kind, int rank, const hsize_t *dims,
                const hsize_t *maxdims, hid_t dcpl, int write) {
  size_t n = 1;
  for (int i = 0; i < rank; ++i) n *= dims[i];
  hid_t file_t, mem_t;
  int own = 0;
  switch (kind) {
    case U8: file_t = H5T_STD_U8LE, mem_t = H5T_NATIVE_UCHAR; break;
    case B8: file_t = H5T_STD_B8LE, mem_t = H5T_NATIVE_B8; break;
    case I32: file_t = H5T_STD_I32LE, mem_t = H5T_NATIVE_INT; break;
    case I32BE: file_t = H5T_STD_I32BE, mem_t = H5T_NATIVE_INT; break;
    case U16: file_t = H5T_STD_U16LE, mem_t = H5T_NATIVE_USHORT; break;
    case F32: file_t = H5T_IEEE_F32LE, mem_t = H5T_NATIVE_FLOAT; break;
    default: file_t = mem_t = bool_type(), own = 1; break;
  }
  hid_t sp = H5Screate_simple(rank, dims, maxdims);
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  CHECK(H5Pset_create_intermediate_group(lcpl, 1));
  hid_t d = H5Dcreate2(loc, name, file_t, sp, lcpl, dcpl, H5P_DEFAULT);
  CHECK(d);
  if (write) {
    void *b = values(kind, n);
    CHECK(H5Dwrite(d, mem_t, H5S_ALL, H5S_ALL, H5P_DEFAULT, b));
    free(b);
  }
  CHECK(H5Dclose(d));
  CHECK(H5Pclose(lcpl));
  CHECK(H5Sclose(sp));
  if (own) H5Tclose(file_t);
}

static hid_t dcpl_chunked(int rank, const hsize_t *chunk, int deflate, int shuffle,
                          int fletcher) {
  hid_t p = H5Pcreate(H5P_DATASET_CREATE);
  CHECK(H5Pset_chunk(p, rank, chunk));
  if (fletcher) CHECK(H5Pset_fletcher32(p));
  if (shuffle) CHECK(H5Pset_shuffle(p));
  if (deflate) CHECK(H5Pset_deflate(p, deflate));
  return p;
}

static void write_files(const char *dir) {
  char path[1024];
  const hsize_t d3[3] = {6, 7, 9}, c3[3] = {2, 3, 4};
  /* 1. default (earliest) format: contiguous, compact, groups as symbol tables */
  snprintf(path, sizeof(path), "%s/v0_contiguous.h5", dir);
  hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  CHECK(f);
  put(f, "gold", U8, 3, d3, NULL, H5P_DEFAULT, 1);
  put(f, "masks/deeper/i32be", I32BE, 3, d3, NULL, H5P_DEFAULT, 1);
  put(f, "f32", F32, 3, d3, NULL, H5P_DEFAULT, 1);
  put(f, "u16", U16, 3, d3, NULL, H5P_DEFAULT, 1);
  put(f, "b8", B8, 2, d3, NULL, H5P_DEFAULT, 1);
  {
    hid_t p = H5Pcreate(H5P_DATASET_CREATE);
    CHECK(H5Pset_layout(p, H5D_COMPACT));
    put(f, "compact_u8", U8, 3, d3, NULL, p, 1);
    H5Pclose(p);
  }
  put(f, "unwritten", U8, 3, d3, NULL, H5P_DEFAULT, 0);
  for (int k = 0; k < 40; ++k) {  /* enough names to split the group B-tree */
    char nm[32];
    snprintf(nm, sizeof(nm), "many/d%02d", k);
    const hsize_t one[1] = {4};
    put(f, nm, U8, 1, one, NULL, H5P_DEFAULT, 1);
  }
  CHECK(H5Fclose(f));

  /* 2. default format, chunked (B-tree v1 index) with filters */
  snprintf(path, sizeof(path), "%s/v0_chunked.h5", dir);
  f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  hid_t p = dcpl_chunked(3, c3, 0, 0, 0);
  put(f, "plain_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 6, 0, 0);
  put(f, "deflate_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 4, 1, 0);
  put(f, "shuffle_deflate_i32", I32, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 1, 0, 1);
  put(f, "fletcher_deflate_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 0, 0, 0);
  put(f, "unwritten_chunked", U8, 3, d3, NULL, p, 0);
  H5Pclose(p);
  {  /* enough chunks for a B-tree of more than one level */
    const hsize_t big[1] = {6000}, cb[1] = {7};
    p = dcpl_chunked(1, cb, 3, 0, 0);
    put(f, "many_chunks_u8", U8, 1, big, NULL, p, 1);
    H5Pclose(p);
  }
  CHECK(H5Fclose(f));

  /* 3. the reference's benchmark layout: two flat bool datasets of one file */
  snprintf(path, sizeof(path), "%s/bool_flat.h5", dir);
  f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  const hsize_t flat[1] = {40 * 30 * 20};
  put(f, "onlyBladderBoolFlat", BOOL_GOLD, 1, flat, NULL, H5P_DEFAULT, 1);
  put(f, "onlyLungsBoolFlat", BOOL_SEGM, 1, flat, NULL, H5P_DEFAULT, 1);
  CHECK(H5Fclose(f));

  /* 4. latest format: superblock 3, object header 2, layout 4 chunk indexes;
   * groups of at most 8 links keep them in the object header (compact) */
  snprintf(path, sizeof(path), "%s/latest.h5", dir);
  hid_t fapl = H5Pcreate(H5P_FILE_ACCESS);
  CHECK(H5Pset_libver_bounds(fapl, H5F_LIBVER_LATEST, H5F_LIBVER_LATEST));
  f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, fapl);
  put(f, "a/contig_u8", U8, 3, d3, NULL, H5P_DEFAULT, 1);
  put(f, "b/sub/i32be", I32BE, 3, d3, NULL, H5P_DEFAULT, 1);
  p = dcpl_chunked(3, c3, 6, 0, 0);
  put(f, "a/fixed_array_deflate_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 0, 0, 0);
  put(f, "a/fixed_array_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, c3, 0, 0, 0);
  CHECK(H5Pset_alloc_time(p, H5D_ALLOC_TIME_EARLY));
  put(f, "a/implicit_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, d3, 5, 1, 0);
  put(f, "a/single_chunk_shuffle_deflate_i32", I32, 3, d3, NULL, p, 1);
  H5Pclose(p);
  p = dcpl_chunked(3, d3, 0, 0, 0);
  put(f, "a/single_chunk_u8", U8, 3, d3, NULL, p, 1);
  H5Pclose(p);
  {  /* more chunks than one fixed-array page (1024) */
    const hsize_t big[1] = {4100}, cb[1] = {2};
    p = dcpl_chunked(1, cb, 0, 0, 0);
    put(f, "b/paged_u8", U8, 1, big, NULL, p, 1);
    H5Pclose(p);
    p = dcpl_chunked(1, cb, 2, 0, 0);
    put(f, "b/paged_deflate_u8", U8, 1, big, NULL, p, 1);
    H5Pclose(p);
  }
  {  /* unlimited extent: extensible-array index (not read) */
    const hsize_t one[1] = {64}, mx[1] = {H5S_UNLIMITED}, cb[1] = {16};
    p = dcpl_chunked(1, cb, 0, 0, 0);
    put(f, "b/unlimited_u8", U8, 1, one, mx, p, 1);
    H5Pclose(p);
  }
  for (int k = 0; k < 12; ++k) {  /* more links than compact storage holds */
    char nm[32];
    snprintf(nm, sizeof(nm), "dense/d%02d", k);
    const hsize_t one[1] = {4};
    put(f, nm, U8, 1, one, NULL, H5P_DEFAULT, 1);
  }
  CHECK(H5Fclose(f));
  H5Pclose(fapl);

  /* 5. a 512-byte user block before the superblock (as MATLAB v7.3 files) */
  snprintf(path, sizeof(path), "%s/userblock.h5", dir);
  hid_t fcpl = H5Pcreate(H5P_FILE_CREATE);
  CHECK(H5Pset_userblock(fcpl, 512));
  f = H5Fcreate(path, H5F_ACC_TRUNC, fcpl, H5P_DEFAULT);
  put(f, "gold", U8, 3, d3, NULL, H5P_DEFAULT, 1);
  CHECK(H5Fclose(f));
  H5Pclose(fcpl);
}

int main(int argc, char **argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: make_h5 <output dir>\n");
    return 2;
  }
  write_files(argv[1]);
  return 0;
}
