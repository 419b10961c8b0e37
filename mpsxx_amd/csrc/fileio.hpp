// On-disk block-sparse arrays and the site-file names of the sweep drivers.
//
// The reference stores each QSTArray through a boost binary archive (src/fileio.h:9-25) under the names of
// mpsxx::getfile (src/fileio.cpp:7-13).  Those archives depend on the boost build and the in-memory layout of the
// BTAS classes, so this library defines its own self-describing format (all fields little-endian):
//
//   char     magic[8]      "MPSXXQST"
//   uint32   version       1
//   uint32   reserved      0
//   int32    rank          0 for an array with no structure (size() == 0), else 1..kMaxRank
//   int32    qsym          1 QSTArray, 0 STArray
//   int32    q_total[2]    particle number, 2Sz
//   int32    nsec[rank]
//   int32    sec_q[S][2]   S = sum(nsec), index 0's sectors first
//   int32    sec_dim[S]
//   int64    nblocks
//   int32    block_idx[nblocks][rank]   strictly ascending block tag (row-major sector index)
//   int64    nelem         sum of the blocks' dense sizes
//   float64  data[nelem]   blocks in the order above, each row-major
//   uint64   checksum      FNV-1a (64-bit) over every preceding byte
//
// A reader rejects a wrong magic or version, sizes out of range, blocks out of order or not allowed by the quantum
// numbers, a short or over-long file and a checksum mismatch.
#pragma once

#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "engine.hpp"

namespace mpx {

struct QstFile {
  bool empty = true;  // rank 0 on disk
  Layout L;           // finalized (canonical offsets)
  std::vector<double> data;
};

// Write L's structure and `data` (L.nelem values in canonical order); L == nullptr writes an empty array.  The file
// is written beside `path` and renamed into place, so a reader never sees a partial file under `path`.
void write_qst(const std::string& path, const Layout* L, const double* data);
// Read and validate a file written by write_qst; with want_data false only the structure is read (the checksum is
// still verified over the whole file).
QstFile read_qst(const std::string& path, bool want_data = true);

// "[prefix]/[name].[site].[istate].[jstate].0.tmp": mpsxx::getfile (src/fileio.cpp:7-13) with the process rank
// field fixed at 0 (one process drives one device).
std::string getfile(const std::string& name, const std::string& prefix, int site, int istate, int jstate);

// Device tensors: download + write_qst, and read_qst + allocate + upload (the tensor's previous contents are
// replaced).  Both synchronize `s`.
void tensor_save(const Tensor& t, const std::string& path, hipStream_t s);
void tensor_load(const std::string& path, Tensor& t, hipStream_t s);

// Small fixed record files (the sweep driver's state record): magic, then int32 fields, then the checksum.
void write_record(const std::string& path, const char (&magic)[9], const std::vector<int32_t>& fields);
std::vector<int32_t> read_record(const std::string& path, const char (&magic)[9], size_t nfields);

}  // namespace mpx
