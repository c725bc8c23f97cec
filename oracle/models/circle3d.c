/*
 * circle3d.c -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * Hand restatement of what the reference C backend emits for
 * examples/circle3d_clustered.abl: examples/circle3d.abl (reference) with a
 * clustered initialisation (num_clusters Gaussian clusters of width
 * cluster_sigma, Box-Muller deviates, plus a uniform background).  With
 * num_clusters = 0 every agent comes from random(environment.max), which is
 * exactly the reference model's initialisation, so this restatement also
 * stands for examples/circle3d.abl.
 *
 * Codegen rules: src/backend/CPrinter.cpp:148-173, 250-302;
 * GenericCPrinter.cpp:20-94; AnalysisVisitor.cpp:89-112, 971-974, 1108-1120
 * (environment.max replaced by its folded value).
 */
#include ABL_RT_HEADER
#include "../orc_near.h"
#include "../orc_model.h"

typedef struct {
  float3 pos;
} Point;
static const type_info Point_info[] = {
  { TYPE_FLOAT3, offsetof(Point, pos), "pos", true },
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
int num_clusters = 64;
abl_float cluster_sigma = 20.0;
abl_float background_fraction = 0.2;
abl_float W = 27.1442;

static orc_grid grid_Point;

float3 random_2(float3 min, float3 max) {
  return float3_create(random_float(min.x, max.x), random_float(min.y, max.y),
                       random_float(min.z, max.z));
}
abl_float random_3(abl_float max) {
  return random_float(0, max);
}
float3 random_5(float3 max) {
  return random_2(float3_fill(0), max);
}
float3 clamp_2(float3 pos, float3 min, float3 max) {
  return float3_create(((pos.x < min.x) ? min.x : ((pos.x > max.x) ? max.x : pos.x)),
                       ((pos.y < min.y) ? min.y : ((pos.y > max.y) ? max.y : pos.y)),
                       ((pos.z < min.z) ? min.z : ((pos.z > max.z) ? max.z : pos.z)));
}
float3 clamp_5(float3 pos, float3 max) {
  return clamp_2(pos, float3_fill(0), max);
}

abl_float gaussian(abl_float sigma) {
  abl_float u1 = random_3(1.0);
  abl_float u2 = random_3(1.0);
  return ((sigma * sqrt((((abl_float) (-2)) * log((1.0 - u1))))) * cos(((2.0 * PI) * u2)));
}

void move_point(Point* in, Point* out) {
  float3 new_pos = in->pos;
  ORC_NEAR_OPEN3(Point, nx, agents.agents_Point, grid_Point, in->pos)
    if (dist_float3(nx->pos, in->pos) > (2.0 * r)) continue;
    ORC_NEAR_HIT(nx, agents.agents_Point, Point)
    {
      abl_float pos_dist = dist_float3(in->pos, nx->pos);
      if ((pos_dist == 0)) continue;
      abl_float sep_dist = (pos_dist - r);
      float3 force = float3_fill(0);
      if ((sep_dist > 0.0)) {
        force = float3_div_scalar(float3_mul_scalar(float3_sub(nx->pos, in->pos), (k_att * ((2.0 * r) - pos_dist))), pos_dist);
      } else {
        force = float3_mul_scalar(float3_sub(in->pos, nx->pos), k_rep);
      }
      new_pos = float3_add(new_pos, force);
    }
  ORC_NEAR_CLOSE
  out->pos = clamp_2(new_pos, float3_fill(1), float3_fill((W - 1.0)));
}

static void model_init(void) {
  int per_cluster = ((num_clusters > 0) ? (((int) (((1.0 - background_fraction) * ((abl_float) num_agents)))) / num_clusters) : 0);
  int num_background = (num_agents - (per_cluster * num_clusters));
  for (int c = 0, _var1 = num_clusters; c < _var1; ++c) {
    float3 centre = random_5(float3_create(W, W, W));
    for (int i = 0, _var2 = per_cluster; i < _var2; ++i) {
      abl_float dx = gaussian(cluster_sigma);
      abl_float dy = gaussian(cluster_sigma);
      abl_float dz = gaussian(cluster_sigma);
      *DYN_ARRAY_PLACE(&agents.agents_Point, Point) = (Point) {
        .pos = clamp_5(float3_add(centre, float3_create(dx, dy, dz)), float3_create(W, W, W)),
      };
    }
  }
  for (int i = 0, _var3 = num_background; i < _var3; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Point, Point) = (Point) {
      .pos = random_5(float3_create(W, W, W)),
    };
  }
}

static double radius_move_point(void) { return (2.0 * r); }

static void model_env(double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = hi[2] = (double) W;
  *g = 2.0 * (double) r;
}

static const orc_param params[] = {
  { "num_timesteps", 'i', &num_timesteps },
  { "num_agents", 'i', &num_agents },
  { "rho", 'd', &rho },
  { "k_rep", 'd', &k_rep },
  { "k_att", 'd', &k_att },
  { "r", 'd', &r },
  { "num_clusters", 'i', &num_clusters },
  { "cluster_sigma", 'd', &cluster_sigma },
  { "background_fraction", 'd', &background_fraction },
  { "W", 'd', &W },
  { NULL, 0, NULL }
};
static const orc_agent agent_table[] = {
  { "Point", Point_info, &agents.agents_Point, &agents.agents_Point_dbuf, sizeof(Point),
    offsetof(Point, pos), 3, &grid_Point },
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
