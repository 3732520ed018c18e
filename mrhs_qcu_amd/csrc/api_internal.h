// Helpers shared by the C ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dslash_dispatch.h"
#include "qcu_mrhs.h"

namespace qmrhs {

void set_last_error(const std::string &m);
int api_fail(const std::string &m);
bool geom_from_c(const qcuMrhsGeom *g, Geom *out, std::string *err);
int prec_bytes(int prec);
long long sites_per_parity(const Geom &g);

Tile tile_for(const Geom &g, int prec, int k);
Tile tile_from_c(const qcuMrhsTile &c);
qcuMrhsTile tile_to_c(const Tile &t);
void set_tile(const Geom &g, int prec, int k, const Tile &t);
bool tuned(const Geom &g, int prec, int k);
std::vector<Tile> tile_candidates(const Geom &g, int prec, int k);

int run_packed_dslash(void *out, const void *in, const void *x, double xa, double xb, const void *gauge, int prec,
                      int k, int parity, int dagger, const Geom &g, const Tile *tile, hipStream_t st,
                      std::string *err);

}  // namespace qmrhs
