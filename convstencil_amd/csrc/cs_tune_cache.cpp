// cs_tune_cache.cpp -- the autotuner's cache: variant choices per (shape, size, dtype).
//
// The reference has no autotuner (its tile shapes and fusion are fixed at compile
// time).  Here a plan created with cs_options.autotune times every compiled variant
// that fits it and keeps the fastest; the choice is cached in memory and in a text
// file, one line per key ("<key> <variant id> <ms per step>", later lines win):
//   $CONVSTENCIL_TUNE_CACHE if set (an empty value keeps the cache in memory only),
//   else $XDG_CACHE_HOME/convstencil_amd/tune.txt, else $HOME/.cache/convstencil_amd/tune.txt.
#include <sys/stat.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

#include "../../include/convstencil.h"

namespace {

struct Entry {
  int32_t variant;
  double ms;
};

struct Cache {
  std::mutex mu;
  bool loaded = false;
  std::string path;  // empty: memory only
  std::map<std::string, Entry> entries;
};

Cache &cache() {
  static Cache c;
  return c;
}

std::string default_path() {
  const char *env = getenv("CONVSTENCIL_TUNE_CACHE");
  if (env) return std::string(env);
  const char *xdg = getenv("XDG_CACHE_HOME");
  if (xdg && *xdg) return std::string(xdg) + "/convstencil_amd/tune.txt";
  const char *home = getenv("HOME");
  if (home && *home) return std::string(home) + "/.cache/convstencil_amd/tune.txt";
  return std::string();
}

// Creates every missing directory of `file`'s parent path.
void make_parents(const std::string &file) {
  for (size_t i = 1; i < file.size(); ++i)
    if (file[i] == '/') {
      const std::string dir = file.substr(0, i);
      if (mkdir(dir.c_str(), 0755) != 0 && errno != EEXIST) return;
    }
}

// Loads the file once (caller holds the lock).
void load_locked(Cache &c) {
  if (c.loaded) return;
  c.loaded = true;
  c.path = default_path();
  if (c.path.empty()) return;
  FILE *f = fopen(c.path.c_str(), "r");
  if (!f) return;
  char line[512];
  while (fgets(line, sizeof line, f)) {
    char key[400];
    int variant = 0;
    double ms = 0.0;
    if (sscanf(line, "%399s %d %lf", key, &variant, &ms) == 3 && variant > 0) c.entries[key] = Entry{variant, ms};
  }
  fclose(f);
}

bool key_ok(const char *key) {
  if (!key || !*key || strlen(key) >= 400) return false;
  for (const char *p = key; *p; ++p)
    if (*p == ' ' || *p == '\n' || *p == '\t') return false;
  return true;
}

}  // namespace

extern "C" {

const char *cs_tune_cache_path(void) {
  Cache &c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  load_locked(c);
  return c.path.c_str();
}

cs_status cs_tune_cache_reset(void) {
  Cache &c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  c.loaded = false;
  c.entries.clear();
  c.path.clear();
  return CS_OK;
}

cs_status cs_tune_lookup(const char *key, int32_t *variant, double *ms_per_step) {
  if (!key_ok(key) || !variant) return CS_ERR_INVALID;
  Cache &c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  load_locked(c);
  auto it = c.entries.find(key);
  if (it == c.entries.end()) return CS_ERR_UNSUPPORTED;
  *variant = it->second.variant;
  if (ms_per_step) *ms_per_step = it->second.ms;
  return CS_OK;
}

cs_status cs_tune_store(const char *key, int32_t variant, double ms_per_step) {
  if (!key_ok(key) || variant <= 0) return CS_ERR_INVALID;
  Cache &c = cache();
  std::lock_guard<std::mutex> g(c.mu);
  load_locked(c);
  c.entries[key] = Entry{variant, ms_per_step};
  if (c.path.empty()) return CS_OK;
  make_parents(c.path);
  FILE *f = fopen(c.path.c_str(), "a");
  if (!f) return CS_OK;  // an unwritable cache file only loses persistence
  fprintf(f, "%s %d %.6g\n", key, variant, ms_per_step);
  fclose(f);
  return CS_OK;
}

}  // extern "C"
