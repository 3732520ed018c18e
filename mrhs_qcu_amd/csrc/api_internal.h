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
// Tile of the user-layout (AoS pointer) path: the tile table's entry for
// precision code prec | QCU_MRHS_TILE_USER_LAYOUT, else a tile without the
// hypercube bit (the pointer launch then keeps its own traversal).
Tile user_layout_tile(const Geom &g, int prec, int k);
Tile tile_from_c(const qcuMrhsTile &c);
qcuMrhsTile tile_to_c(const Tile &t);
void set_tile(const Geom &g, int prec, int k, const Tile &t);
bool tuned(const Geom &g, int prec, int k);
std::vector<Tile> tile_candidates(const Geom &g, int prec, int k);

int run_packed_dslash(void *out, const void *in, const void *x, double xa, double xb, const void *gauge, int prec,
                      int k, int parity, int dagger, const Geom &g, const Tile *tile, hipStream_t st,
                      std::string *err);

// qcuMrhsDslashEx in three steps, shared with its allocation-free variant:
// the call (tuned tile for the shape), the bytes of norm partials it needs
// (per-workgroup partials and the reduction's stage buffer), and the launch
// plus the per-RHS reduction into norm2_out (when not null) using `scratch`
// of at least that size. Asynchronous on the call's stream.
bool dslash_ex_call(DslashCall *c, void *out, int out_prec, const void *in, const void *x, int in_prec, double a,
                    double b, const float *sc_in, const float *sc_x, const float *sc_out, const void *gauge, int nrhs,
                    int parity, int dagger, const qcuMrhsGeom *gc, void *stream, std::string *err);
size_t dslash_ex_scratch_bytes(const DslashCall &c);
int dslash_ex_launch(const DslashCall &c, double *norm2_out, double *scratch, std::string *err);

}  // namespace qmrhs
