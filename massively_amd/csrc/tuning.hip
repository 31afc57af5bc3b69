// Autotuned choices (no counterpart in the reference, whose kernels are
// specialised by its JIT compiler): one process-wide table of
// (device signature, family, key) -> chosen candidate, which can be saved to
// and loaded from a text file so that a later process starts tuned.
//
// File format, one entry per line (lines starting with '#' are comments):
//   <device signature> <family> <key> <choice>
// The device signature is "<gfx arch name>/<compute units>", so entries made
// on one kind of GPU are never used on another.
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "msv_internal.hpp"

namespace msv {

namespace {

using TuneKey = std::tuple<std::string, std::string, uint32_t>;

std::mutex g_mu;
std::map<TuneKey, int> g_table;
bool g_env_loaded = false;

// Merges the entries of `path` into the table (g_mu held). False when the
// file cannot be read or holds a malformed line.
bool load_locked(const char* path) {
  FILE* f = fopen(path, "r");
  if (!f) return false;
  char line[512];
  bool ok = true;
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char sig[128], fam[64];
    unsigned key = 0;
    int choice = 0;
    if (sscanf(line, "%127s %63s %u %d", sig, fam, &key, &choice) != 4) {
      ok = false;
      continue;
    }
    g_table[TuneKey(sig, fam, key)] = choice;
  }
  fclose(f);
  return ok;
}

// Writes the table to `path` through a temporary file and a rename (g_mu
// held).
bool save_locked(const char* path) {
  const std::string tmp = std::string(path) + ".tmp";
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return false;
  fprintf(f, "# massively_amd tuning cache: <device signature> <family> <key> <choice>\n");
  for (const auto& kv : g_table)
    fprintf(f, "%s %s %u %d\n", std::get<0>(kv.first).c_str(), std::get<1>(kv.first).c_str(),
            std::get<2>(kv.first), kv.second);
  const bool ok = fclose(f) == 0;
  return ok && rename(tmp.c_str(), path) == 0;
}

const char* env_path() {
  const char* p = getenv("MSV_TUNING_CACHE");
  return p && *p ? p : nullptr;
}

void env_load_locked() {
  if (g_env_loaded) return;
  g_env_loaded = true;
  if (const char* p = env_path()) load_locked(p);
}

int get_locked(const std::string& sig, const char* family, uint32_t key) {
  env_load_locked();
  auto it = g_table.find(TuneKey(sig, family, key));
  return it == g_table.end() ? -1 : it->second;
}

}  // namespace

int tune_get(const msv_executor* ex, const char* family, uint32_t key) {
  std::lock_guard<std::mutex> g(g_mu);
  return get_locked(ex->signature, family, key);
}

void tune_put(const msv_executor* ex, const char* family, uint32_t key, int choice) {
  std::lock_guard<std::mutex> g(g_mu);
  env_load_locked();
  g_table[TuneKey(ex->signature, family, key)] = choice;
  if (const char* p = env_path()) save_locked(p);
}

int tune_get_current(const char* family, uint32_t key) {
  int dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  char sig[96];
  snprintf(sig, sizeof sig, "%s/%d", prop.gcnArchName, prop.multiProcessorCount);
  for (char* q = sig; *q; ++q)
    if (*q == ' ') *q = '_';
  std::lock_guard<std::mutex> g(g_mu);
  return get_locked(sig, family, key);
}

}  // namespace msv

using namespace msv;

extern "C" {

msv_status msv_tuning_save(const char* path) {
  clear_error();
  if (!path) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null path");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (!save_locked(path)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "cannot write tuning file %s", path);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return MSV_OK;
}

msv_status msv_tuning_load(const char* path) {
  clear_error();
  if (!path) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "null path");
    return MSV_ERR_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> g(g_mu);
  if (!load_locked(path)) {
    set_error(MSV_ERR_INVALID_ARGUMENT, 0, 0, "cannot read tuning file %s (or a malformed line)",
              path);
    return MSV_ERR_INVALID_ARGUMENT;
  }
  return MSV_OK;
}

msv_status msv_tuning_clear(void) {
  clear_error();
  std::lock_guard<std::mutex> g(g_mu);
  g_table.clear();
  return MSV_OK;
}

uint32_t msv_tuning_count(void) {
  std::lock_guard<std::mutex> g(g_mu);
  return (uint32_t)g_table.size();
}

}  // extern "C"
