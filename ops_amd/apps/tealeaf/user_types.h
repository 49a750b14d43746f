// TeaLeaf 2D: types shared by the host driver and the device kernels. The state, grid and
// field records keep the layout of the reference application
// (apps/c/TeaLeaf/user_types.h) since they are uploaded as ops_decl_const data.
#ifndef TEA_USER_TYPES_H
#define TEA_USER_TYPES_H

#define FIELD_DENSITY 0
#define FIELD_ENERGY0 1
#define FIELD_ENERGY1 2
#define FIELD_U 3
#define FIELD_P 4
#define FIELD_SD 5
#define NUM_FIELDS 6

#define CONDUCTIVITY 1
#define RECIP_CONDUCTIVITY 2

typedef struct {
  int defined;
  double density, energy;
  int geometry;
  double xmin, xmax, ymin, ymax, radius;
} state_type;

typedef struct grid_type {
  double xmin, ymin, xmax, ymax;
  int x_cells, y_cells;
} grid_type;

typedef struct field_type {
  int left, right, bottom, top, left_boundary, right_boundary, bottom_boundary, top_boundary;
  int x_min, y_min, x_max, y_max;
} field_type;

#endif
