// Lattice and process-grid descriptors (interface of the reference's
// include/qcuDesc/qcu_desc.h:9-97: same members, constructors and accessors).
#pragma once

#include <cassert>
#include <cstdio>

#include "qcu.h"
#include "qcu_enum.h"

template <int _Ndim>
struct QcuLattDesc {
  int dim[_Ndim];

  QcuLattDesc(int x, int y, int z, int t) { set(x, y, z, t); }
  QcuLattDesc(QcuParam *param)
  {
    set(param->lattice_size[X_DIM], param->lattice_size[Y_DIM], param->lattice_size[Z_DIM],
        param->lattice_size[T_DIM]);
  }

  int X() { return dim[X_DIM]; }
  int Y() { return get(Y_DIM); }
  int Z() { return get(Z_DIM); }
  int T() { return get(T_DIM); }

 private:
  void set(int x, int y, int z, int t)
  {
    const int v[4] = {x, y, z, t};
    for (int d = 0; d < _Ndim && d < 4; ++d) dim[d] = v[d];
  }
  int get(int d)
  {
    if (d < _Ndim) return dim[d];
    fprintf(stderr, "file %s line %d dim = %d\n", __FILE__, __LINE__, _Ndim);
    assert(0);
    return 0;
  }
};

template <int _Ndim>
struct QcuProcDesc {
  int dim[_Ndim];

  QcuProcDesc(int x, int y, int z, int t) { set(x, y, z, t); }
  QcuProcDesc(QcuGrid *grid)
  {
    set(grid->grid_size[X_DIM], grid->grid_size[Y_DIM], grid->grid_size[Z_DIM], grid->grid_size[T_DIM]);
  }

  int X() { return dim[X_DIM]; }
  int Y() { return get(Y_DIM); }
  int Z() { return get(Z_DIM); }
  int T() { return get(T_DIM); }

 private:
  void set(int x, int y, int z, int t)
  {
    const int v[4] = {x, y, z, t};
    for (int d = 0; d < _Ndim && d < 4; ++d) dim[d] = v[d];
  }
  int get(int d)
  {
    if (d < _Ndim) return dim[d];
    fprintf(stderr, "file %s line %d dim = %d\n", __FILE__, __LINE__, _Ndim);
    assert(0);
    return 0;
  }
};
