// cli.cpp — mlem_cli, the drop-in for the reference's sequential driver:
//
//   mlem_cli <matrix> <input> <output> <iterations> [<checkpointing>]
//
// Same positional arguments, stdout lines and raw fp32 output as
// mlem_seq/seqcsr4mlem.cpp:22-34,192-211; the optional fifth argument is the
// checkpoint interval of the MPI/CUDA drivers (mpicsr4mlem.cpp:165-173,
// hybCalcFunc.cu:852-853; 0 = off). Where the reference lets an exception or
// a failed assert abort the process, this driver prints the message and
// exits with status 1.
//
// Environment: MLEM_MODE = fast (default) | exact | atomic;
// MLEM_SYMMETRY = expand regenerates the elements that a csr4 file with
// symmetry flags omits (default: use the stored elements, as the reference).
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "csr4.hpp"
#include "mlem.h"

namespace {

int fail(const std::string& msg) {
  std::cerr << "mlem_cli: " << msg << std::endl;
  return 1;
}

int parse_int(const char* s, const char* what, int* out) {
  try {
    size_t pos = 0;
    const int v = std::stoi(s, &pos);
    if (pos != std::strlen(s)) throw std::invalid_argument(s);
    *out = v;
    return 0;
  } catch (const std::exception&) {
    std::cerr << "mlem_cli: invalid " << what << ": " << s << std::endl;
    return 1;
  }
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc != 5 && argc != 6) {
    std::cerr << "wrong number of command line parameters\n"
              << "usage: " << argv[0]
              << " <matrix> <input> <output> <iterations> [<checkpointing>]" << std::endl;
    return 1;
  }
  const std::string mtx = argv[1], in = argv[2], out = argv[3];
  int iterations = 0, checkpointing = 0;
  if (parse_int(argv[4], "iteration count", &iterations)) return 1;
  if (argc == 6 && parse_int(argv[5], "checkpoint interval", &checkpointing)) return 1;
  if (iterations < 0) return fail("iteration count must be >= 0");
  if (checkpointing < 0) return fail("checkpoint interval must be >= 0");

  int mode = MLEM_MODE_FAST;
  if (const char* m = std::getenv("MLEM_MODE")) {
    if (!std::strcmp(m, "exact")) mode = MLEM_MODE_EXACT;
    else if (!std::strcmp(m, "atomic")) mode = MLEM_MODE_ATOMIC;
    else if (*m && std::strcmp(m, "fast")) return fail(std::string("unknown MLEM_MODE ") + m);
  }
  unsigned open_flags = mode == MLEM_MODE_ATOMIC ? MLEM_OPEN_NO_CSC : 0u;
  if (const char* s = std::getenv("MLEM_SYMMETRY")) {
    if (!std::strcmp(s, "expand")) open_flags |= MLEM_OPEN_EXPAND_SYMMETRY;
    else if (*s && std::strcmp(s, "stored"))
      return fail(std::string("unknown MLEM_SYMMETRY ") + s);
  }

  // MLEM_STREAM=<MiB>: keep the matrix in host memory and stream it in row
  // blocks of at most that many MiB per part (0: sized from free device memory)
  bool streamed = false;
  uint64_t block_bytes = 0;
  if (const char* s = std::getenv("MLEM_STREAM")) {
    if (*s) {
      int mib = 0;
      if (parse_int(s, "MLEM_STREAM block size (MiB)", &mib)) return 1;
      if (mib < 0) return fail("MLEM_STREAM must be >= 0");
      streamed = true;
      block_bytes = (uint64_t)mib << 20;
    }
  }
  // MLEM_STREAM_RESIDENT=<MiB>: of a streamed matrix, keep that many MiB of
  // blocks in device memory
  uint64_t resident_bytes = 0;
  if (const char* s = std::getenv("MLEM_STREAM_RESIDENT")) {
    if (*s) {
      int mib = 0;
      if (parse_int(s, "MLEM_STREAM_RESIDENT (MiB)", &mib)) return 1;
      if (mib < 0) return fail("MLEM_STREAM_RESIDENT must be >= 0");
      resident_bytes = (uint64_t)mib << 20;
    }
  }

  std::cout << "Matrix file: " << mtx << std::endl;
  std::cout << "Input file: " << in << std::endl;
  std::cout << "Output file: " << out << std::endl;
  std::cout << "Iterations: " << iterations << std::endl;
  if (argc == 6) std::cout << "Checkpointing: " << checkpointing << std::endl;

  std::vector<int32_t> y;
  try {
    y = mlem::read_raw_i32(in);
  } catch (const mlem::Csr4Error& e) {
    return fail(e.msg);
  }

  mlem_matrix* m = nullptr;
  const int orc = streamed ? mlem_open_streamed(mtx.c_str(), open_flags & MLEM_OPEN_EXPAND_SYMMETRY,
                                                 block_bytes, resident_bytes, &m)
                           : mlem_open(mtx.c_str(), open_flags, &m);
  if (orc != MLEM_OK) return fail(mlem_last_error());
  uint32_t rows = 0, cols = 0;
  uint64_t nnz = 0;
  mlem_dims(m, &rows, &cols, &nnz);
  if (y.size() != rows) {
    mlem_close(m);
    return fail("input file " + in + " holds " + std::to_string(y.size()) +
                " counts but the matrix has " + std::to_string(rows) + " rows");
  }
  std::vector<float> image(cols, 0.0f);
  mlem_options o;
  mlem_options_default(&o);
  o.mode = mode;
  o.verbose = 1;
  // The 5-argument form restarts from a complete checkpoint even when its
  // <checkpointing> argument is 0 (mpicsr4mlem.cpp:536-549); it then writes
  // none. The 4-argument form is seqcsr4mlem, which has no checkpoints.
  o.checkpoint_interval = argc == 6 && checkpointing == 0 ? MLEM_CHECKPOINT_RESTORE_ONLY
                                                          : checkpointing;
  int rc = mlem_reconstruct(m, y.data(), rows, image.data(), cols, iterations, &o);
  if (rc != MLEM_OK) {
    const std::string msg = mlem_last_error();
    mlem_close(m);
    return fail(msg);
  }
  mlem_close(m);
  try {
    mlem::write_raw(out, image.data(), image.size() * sizeof(float), false);
  } catch (const mlem::Csr4Error& e) {
    return fail(e.msg);
  }
  return 0;
}
