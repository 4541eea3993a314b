// Measured launch shapes of the occupied-factor J/K build.
//
// The U kernel's i-tile width and the exchange SYRK's split-K slice count are
// chosen by fixed rules (occupied_tile, the occupancy rule in exchange_syrk).
// DfPlan::autotune times alternatives on the plan's own tensors and records
// the fastest here, per device architecture and build shape; plans consult
// the registry when they lay out an occupied factor and launch the SYRK.  The
// registry starts empty (the rules apply), so nothing changes unless a caller
// tunes or loads a profile.  With VIBEQC_TUNING_FILE set, the file is read on
// first use and rewritten after every recorded shape.
//
// File format (text, one entry per line after the header line
// "vibeqc-jk-tuning 1"):
//   <arch> <precision 0|1|2> <nbf> <occupied columns> <ldq> <u_tile> <syrk_splits> <ms>
// where arch is the device's gfx name, u_tile 0..10 (0: the rule) and
// syrk_splits >= 0 (0: the rule).  Entries for other architectures are kept
// and written back but never match.
#pragma once

#include <string>
#include <vector>

namespace vqc {

struct JkShape {
  int u_tile = 0;       // U i-tile in 32-column units (1..10), 0: occupied_tile(r)
  int syrk_splits = 0;  // split-K slices per SYRK run, 0: the occupancy rule
};

struct TuningEntry {
  std::string arch;
  int precision = 0;  // Precision as int (FP64 0, FP32 1, BF16X3 2)
  int n = 0, r = 0, ldq = 0;
  JkShape shape;
  double ms = 0.0;  // measured build time of the recorded shape
};

// The recorded shape for this key, if any.
bool tuned_shape(const std::string& arch, int precision, int n, int r, int ldq, JkShape& out);
// Records (or replaces) an entry; rewrites VIBEQC_TUNING_FILE when it is set.
// Throws std::invalid_argument for a shape out of range.
void record_shape(const TuningEntry& e);
std::vector<TuningEntry> tuned_shapes();
void clear_tuning();
// Reads entries from a file (replacing entries with equal keys); returns false
// when the file does not exist, throws std::invalid_argument when it is malformed.
bool load_tuning(const std::string& path);
void save_tuning(const std::string& path);

}  // namespace vqc
