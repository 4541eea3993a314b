// Registry of measured J/K launch shapes (see tuning.hpp).
#include "tuning.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <stdexcept>
#include <tuple>

namespace vqc {
namespace {

using Key = std::tuple<std::string, int, int, int, int>;

struct Registry {
  std::mutex mutex;
  std::map<Key, TuningEntry> entries;
  bool env_loaded = false;
};

Registry& registry() {
  static Registry r;
  return r;
}

Key key_of(const TuningEntry& e) { return Key(e.arch, e.precision, e.n, e.r, e.ldq); }

void validate(const TuningEntry& e) {
  if (e.arch.empty() || e.arch.find_first_of(" \t\r\n") != std::string::npos)
    throw std::invalid_argument("tuning entry needs a device architecture name without spaces");
  if (e.precision < 0 || e.precision > 2) throw std::invalid_argument("tuning entry precision must be 0, 1 or 2");
  if (e.n <= 0 || e.r <= 0 || e.r > e.n || e.ldq <= 0 || e.ldq % 128 != 0)
    throw std::invalid_argument("tuning entry shape out of range (n > 0, 0 < r <= n, ldq a multiple of 128)");
  if (e.shape.u_tile < 0 || e.shape.u_tile > 10) throw std::invalid_argument("tuning entry u_tile must be 0..10");
  if (e.shape.syrk_splits < 0 || e.shape.syrk_splits > 128)
    throw std::invalid_argument("tuning entry syrk_splits must be 0..128");
}

constexpr const char* kHeader = "vibeqc-jk-tuning 1";

// Caller holds the mutex.
bool load_locked(Registry& reg, const std::string& path) {
  std::ifstream in(path);
  if (!in) return false;
  std::string line;
  if (!std::getline(in, line) || line != kHeader)
    throw std::invalid_argument("tuning file " + path + " lacks the header line \"" + kHeader + "\"");
  std::map<Key, TuningEntry> read;
  int lineno = 1;
  while (std::getline(in, line)) {
    ++lineno;
    if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
    std::istringstream ls(line);
    TuningEntry e;
    std::string extra;
    if (!(ls >> e.arch >> e.precision >> e.n >> e.r >> e.ldq >> e.shape.u_tile >> e.shape.syrk_splits >> e.ms) ||
        (ls >> extra))
      throw std::invalid_argument("tuning file " + path + " line " + std::to_string(lineno) + " is malformed");
    try {
      validate(e);
    } catch (const std::invalid_argument& err) {
      throw std::invalid_argument("tuning file " + path + " line " + std::to_string(lineno) + ": " + err.what());
    }
    read[key_of(e)] = e;
  }
  for (auto& kv : read) reg.entries[kv.first] = kv.second;
  return true;
}

// Caller holds the mutex.
void save_locked(const Registry& reg, const std::string& path) {
  const std::string tmp = path + ".tmp";
  {
    std::ofstream out(tmp, std::ios::trunc);
    if (!out) throw std::runtime_error("cannot write tuning file " + tmp);
    out << kHeader << "\n";
    char buf[64];
    for (const auto& kv : reg.entries) {
      const TuningEntry& e = kv.second;
      std::snprintf(buf, sizeof(buf), "%.6g", e.ms);
      out << e.arch << ' ' << e.precision << ' ' << e.n << ' ' << e.r << ' ' << e.ldq << ' ' << e.shape.u_tile << ' '
          << e.shape.syrk_splits << ' ' << buf << "\n";
    }
    if (!out) throw std::runtime_error("cannot write tuning file " + tmp);
  }
  if (std::rename(tmp.c_str(), path.c_str()) != 0) throw std::runtime_error("cannot replace tuning file " + path);
}

// Caller holds the mutex.
void load_env_once(Registry& reg) {
  if (reg.env_loaded) return;
  reg.env_loaded = true;
  if (const char* path = std::getenv("VIBEQC_TUNING_FILE"))
    if (path[0]) load_locked(reg, path);
}

}  // namespace

bool tuned_shape(const std::string& arch, int precision, int n, int r, int ldq, JkShape& out) {
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  load_env_once(reg);
  auto it = reg.entries.find(Key(arch, precision, n, r, ldq));
  if (it == reg.entries.end()) return false;
  out = it->second.shape;
  return true;
}

void record_shape(const TuningEntry& e) {
  validate(e);
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  load_env_once(reg);
  reg.entries[key_of(e)] = e;
  if (const char* path = std::getenv("VIBEQC_TUNING_FILE"))
    if (path[0]) save_locked(reg, path);
}

std::vector<TuningEntry> tuned_shapes() {
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  load_env_once(reg);
  std::vector<TuningEntry> out;
  for (const auto& kv : reg.entries) out.push_back(kv.second);
  return out;
}

void clear_tuning() {
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  reg.entries.clear();
  reg.env_loaded = true;  // an explicit clear also drops what the environment file would add
}

bool load_tuning(const std::string& path) {
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  load_env_once(reg);
  return load_locked(reg, path);
}

void save_tuning(const std::string& path) {
  Registry& reg = registry();
  std::lock_guard<std::mutex> lock(reg.mutex);
  load_env_once(reg);
  save_locked(reg, path);
}

}  // namespace vqc
