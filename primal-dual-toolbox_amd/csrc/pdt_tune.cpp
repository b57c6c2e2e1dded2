// pdt_tune.cpp -- tile-shape cache (see pdt_tune.hpp).
#include "pdt_tune.hpp"

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <tuple>

namespace pdt {

bool TuneKey::operator<(const TuneKey& o) const {
  return std::tie(kind, dt, W, H, Hk, C, nc, ns) < std::tie(o.kind, o.dt, o.W, o.H, o.Hk, o.C, o.nc, o.ns);
}

std::string TuneKey::str() const {
  std::ostringstream ss;
  ss << kind << ' ' << dt << ' ' << W << ' ' << H << ' ' << Hk << ' ' << C << ' ' << nc << ' ' << ns;
  return ss.str();
}

namespace {
std::mutex g_mu;
std::map<TuneKey, TuneVal> g_cache;
bool g_loaded = false;

const char* cache_path() { return std::getenv("PDT_TUNE_CACHE"); }

// File format: one entry per line,
// "kind dt W H Hk C nc ns : cb cg rpi [cb_a [wpc0 wpc1 wpc2 wpc3 [tbx tby tbk [dpx]]]]"
// (cb_a absent: the adjoint row pass uses cb; occupancy caps absent: none;
// blocked ROF tiles or dual pixels per thread absent: the default).
void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  const char* p = cache_path();
  if (!p) return;
  std::ifstream in(p);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    TuneKey k{};
    TuneVal v{};
    char colon = 0;
    if (ss >> k.kind >> k.dt >> k.W >> k.H >> k.Hk >> k.C >> k.nc >> k.ns >> colon >> v.cb >> v.cg >> v.rpi &&
        colon == ':') {
      if (!(ss >> v.cb_a)) v.cb_a = 0;
      int w[4];
      if (ss >> w[0] >> w[1] >> w[2] >> w[3]) {
        for (int i = 0; i < 4; ++i) v.wpc[i] = w[i];
        int t[3];
        if (ss >> t[0] >> t[1] >> t[2]) {
          for (int i = 0; i < 3; ++i) v.tb[i] = t[i];
          if (!(ss >> v.dual_px)) v.dual_px = 0;
          if (!(ss >> v.unpaired)) v.unpaired = 0;
        }
      }
      g_cache[k] = v;
    }
  }
}

void save_locked() {
  const char* p = cache_path();
  if (!p) return;
  std::ofstream out(p, std::ios::trunc);
  for (const auto& kv : g_cache)
    out << kv.first.str() << " : " << kv.second.cb << ' ' << kv.second.cg << ' ' << kv.second.rpi << ' '
        << kv.second.cb_a << ' ' << kv.second.wpc[0] << ' ' << kv.second.wpc[1] << ' ' << kv.second.wpc[2] << ' '
        << kv.second.wpc[3] << ' ' << kv.second.tb[0] << ' ' << kv.second.tb[1] << ' ' << kv.second.tb[2] << ' ' << kv.second.dual_px << ' '
        << kv.second.unpaired << '\n';
}
}  // namespace

TuneVal tune_lookup(const TuneKey& k) {
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  auto it = g_cache.find(k);
  return it == g_cache.end() ? TuneVal{} : it->second;
}

void tune_store(const TuneKey& k, const TuneVal& v) {
  std::lock_guard<std::mutex> lk(g_mu);
  load_locked();
  g_cache[k] = v;
  save_locked();
}

}  // namespace pdt
