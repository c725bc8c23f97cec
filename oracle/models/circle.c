/*
 * circle.c -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * Hand restatement of what the reference C backend emits for
 * examples/circle.abl (2-D): struct layout and RTTI from
 * src/backend/CPrinter.cpp:250-302, globals from GenericCPrinter.cpp:77-94
 * (folded constants printed at 6 significant digits, GenericPrinter.cpp:135-146),
 * lib.abl helpers with their overload names (src/AnalysisVisitor.cpp:209-213),
 * the near loop of CPrinter.cpp:148-173 and the expression rules of
 * GenericCPrinter.cpp:20-75 (scalar*vector printed as vector*scalar,
 * src/AnalysisVisitor.cpp:971-974; int literals promoted to float literals,
 * src/AnalysisVisitor.cpp:89-112).
 *
 * Globals carry the example's defaults; tests overwrite them through the
 * parameter table exactly as `-P name=value` would re-emit them.
 */
#include ABL_RT_HEADER
#include "../orc_near.h"
#include "../orc_model.h"

typedef struct {
  float2 pos;
} Point;
static const type_info Point_info[] = {
  { TYPE_FLOAT2, offsetof(Point, pos), "pos", true },
  { TYPE_END, sizeof(Point), NULL }
};
struct agent_struct {
  dyn_array agents_Point;
  dyn_array agents_Point_dbuf;
};
struct agent_struct agents;
static const agent_info agents_info[] = {
  { Point_info, offsetof(struct agent_struct, agents_Point), "Point" },
  { NULL, 0, NULL }
};

abl_float PI = 3.14159;
int num_timesteps = 100;
int num_agents = 1000;
abl_float rho = 0.05;
abl_float k_rep = 0.05;
abl_float k_att = 0.01;
abl_float r = 5.0;
abl_float W = 141.421;

static orc_grid grid_Point;

float2 random_1(float2 min, float2 max) {
  return float2_create(random_float(min.x, max.x), random_float(min.y, max.y));
}
float2 random_4(float2 max) {
  return random_1(float2_fill(0), max);
}
float2 clamp_1(float2 pos, float2 min, float2 max) {
  return float2_create(((pos.x < min.x) ? min.x : ((pos.x > max.x) ? max.x : pos.x)),
                       ((pos.y < min.y) ? min.y : ((pos.y > max.y) ? max.y : pos.y)));
}

void move_point(Point* in, Point* out) {
  float2 new_pos = in->pos;
  ORC_NEAR_OPEN(Point, nx, agents.agents_Point, grid_Point, in->pos)
    if (dist_float2(nx->pos, in->pos) > (2.0 * r)) continue;
    ORC_NEAR_HIT(nx, agents.agents_Point, Point)
    {
      abl_float pos_dist = dist_float2(in->pos, nx->pos);
      if ((pos_dist == 0)) continue;
      abl_float sep_dist = (pos_dist - r);
      float2 force = float2_fill(0);
      if ((sep_dist > 0.0)) {
        force = float2_div_scalar(float2_mul_scalar(float2_sub(nx->pos, in->pos), (k_att * ((2.0 * r) - pos_dist))), pos_dist);
      } else {
        force = float2_mul_scalar(float2_sub(in->pos, nx->pos), k_rep);
      }
      new_pos = float2_add(new_pos, force);
    }
  ORC_NEAR_CLOSE
  out->pos = clamp_1(new_pos, float2_fill(1), float2_fill((W - 1.0)));
}

/* main() statements before simulate(...) */
static void model_init(void) {
  for (int i = 0, _var1 = num_agents; i < _var1; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Point, Point) = (Point) {
      .pos = random_4(float2_fill(W)),
    };
  }
}

static double radius_move_point(void) { return (2.0 * r); }

static void model_env(double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = (double) W;
  hi[2] = 0.0;
  *g = 2.0 * (double) r; /* the largest constant near radius */
}

static const orc_param params[] = {
  { "num_timesteps", 'i', &num_timesteps },
  { "num_agents", 'i', &num_agents },
  { "rho", 'd', &rho },
  { "k_rep", 'd', &k_rep },
  { "k_att", 'd', &k_att },
  { "r", 'd', &r },
  { "W", 'd', &W },
  { NULL, 0, NULL }
};
static const orc_agent agent_table[] = {
  { "Point", Point_info, &agents.agents_Point, &agents.agents_Point_dbuf, sizeof(Point),
    offsetof(Point, pos), 2, &grid_Point },
};
static const orc_step step_table[] = {
  { "move_point", 0, (void (*)(void *, void *)) move_point, 0, radius_move_point },
};
const orc_model_desc orc_model = {
  params, agent_table, 1, step_table, 1, model_env, model_init, agents_info, &agents,
};

#ifdef ORACLE_MAIN
#include "../orc_main.h"
ORC_DEFINE_MAIN("points.json")
#endif
