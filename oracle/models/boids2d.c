/*
 * boids2d.c -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * Hand restatement of what the reference C backend emits for
 * examples/boids2d_cfg.abl: examples/boids2d.abl (reference) with its flocking
 * weights, time scales and speed limit turned into `param`s.  The speed clamp
 * is written `v / s * max_speed`; with max_speed = 1 that is bit-identical to
 * the reference model's `v /= s` (multiplying by 1.0 is exact), so the same
 * restatement also stands for examples/boids2d.abl with
 * steer/collision/match = 0.65/0.75/1.25 and max_speed = 1.
 *
 * Codegen rules followed: src/backend/CPrinter.cpp:148-173 (near loop),
 * 250-302 (structs and RTTI), 84-90 (agent creation); GenericCPrinter.cpp:20-75
 * (vector operators); AnalysisVisitor.cpp:89-112, 971-974 (promotion and
 * scalar*vector normalisation).
 */
#include ABL_RT_HEADER
#include "../orc_near.h"
#include "../orc_model.h"

typedef struct {
  float2 pos;
  float2 velocity;
} Boid;
static const type_info Boid_info[] = {
  { TYPE_FLOAT2, offsetof(Boid, pos), "pos", true },
  { TYPE_FLOAT2, offsetof(Boid, velocity), "velocity", false },
  { TYPE_END, sizeof(Boid), NULL }
};
struct agent_struct {
  dyn_array agents_Boid;
  dyn_array agents_Boid_dbuf;
};
struct agent_struct agents;
static const agent_info agents_info[] = {
  { Boid_info, offsetof(struct agent_struct, agents_Boid), "Boid" },
  { NULL, 0, NULL }
};

abl_float PI = 3.14159;
abl_float interaction_radius = 10.0;
abl_float separation_radius = 2.0;
abl_float agent_density = 0.065536;
abl_float time_scale = 0.0005;
abl_float global_scale = 0.15;
abl_float steer_scale = 1.0;
abl_float collision_scale = 1.5;
abl_float match_scale = 1.0;
abl_float max_speed = 2.0;
int num_agents = 65536;
int num_timesteps = 200;
abl_float min_pos = 0.0;
abl_float max_pos = 1000.0;

static orc_grid grid_Boid;

float2 random_1(float2 min, float2 max) {
  return float2_create(random_float(min.x, max.x), random_float(min.y, max.y));
}

float2 boundPosition(float2 pos) {
  return float2_create(((pos.x < min_pos) ? max_pos : ((pos.x > max_pos) ? min_pos : pos.x)),
                       ((pos.y < min_pos) ? max_pos : ((pos.y > max_pos) ? min_pos : pos.y)));
}

void update_boid(Boid* in, Boid* out) {
  float2 global_velocity = float2_create(0, 0);
  float2 global_center = float2_create(0, 0);
  float2 collision_center = float2_create(0, 0);
  int interaction_count = 0;
  int collision_count = 0;
  ORC_NEAR_OPEN(Boid, nx, agents.agents_Boid, grid_Boid, in->pos)
    if (dist_float2(nx->pos, in->pos) > interaction_radius) continue;
    ORC_NEAR_HIT(nx, agents.agents_Boid, Boid)
    {
      global_center = float2_add(global_center, nx->pos);
      global_velocity = float2_add(global_velocity, nx->velocity);
      interaction_count += 1;
      abl_float separation = dist_float2(in->pos, nx->pos);
      if ((separation < separation_radius)) {
        collision_center = float2_add(collision_center, nx->pos);
        collision_count += 1;
      }
    }
  ORC_NEAR_CLOSE
  float2 velocity_change = float2_create(0, 0);
  float2 steer_velocity = float2_create(0, 0);
  if ((interaction_count > 0)) {
    global_center = float2_div_scalar(global_center, interaction_count);
    steer_velocity = float2_mul_scalar(float2_sub(global_center, in->pos), steer_scale);
  }
  velocity_change = float2_add(velocity_change, steer_velocity);
  float2 match_velocity = float2_create(0, 0);
  if ((interaction_count > 0)) {
    global_velocity = float2_div_scalar(global_velocity, interaction_count);
    match_velocity = float2_mul_scalar(global_velocity, match_scale);
  }
  velocity_change = float2_add(velocity_change, match_velocity);
  float2 avoid_velocity = float2_create(0, 0);
  if ((collision_count > 0)) {
    collision_center = float2_div_scalar(collision_center, collision_count);
    avoid_velocity = float2_mul_scalar(float2_sub(in->pos, collision_center), collision_scale);
  }
  velocity_change = float2_add(velocity_change, avoid_velocity);
  float2 new_velocity = float2_add(in->velocity, float2_mul_scalar(velocity_change, global_scale));
  abl_float new_velocity_scale = length_float2(new_velocity);
  if ((new_velocity_scale > max_speed)) {
    new_velocity = float2_mul_scalar(float2_div_scalar(new_velocity, new_velocity_scale), max_speed);
  }
  float2 new_pos = float2_add(in->pos, float2_mul_scalar(new_velocity, time_scale));
  new_pos = boundPosition(new_pos);
  out->pos = new_pos;
  out->velocity = new_velocity;
}

static void model_init(void) {
  for (int i = 0, _var1 = num_agents; i < _var1; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Boid, Boid) = (Boid) {
      .pos = random_1(float2_fill(min_pos), float2_fill(max_pos)),
      .velocity = random_1(float2_fill((-1)), float2_fill(1)),
    };
  }
}

static double radius_update_boid(void) { return (double) interaction_radius; }

static void model_env(double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = (double) min_pos;
  hi[0] = hi[1] = (double) max_pos;
  lo[2] = hi[2] = 0.0;
  *g = (double) interaction_radius;
}

static const orc_param params[] = {
  { "interaction_radius", 'd', &interaction_radius },
  { "separation_radius", 'd', &separation_radius },
  { "agent_density", 'd', &agent_density },
  { "time_scale", 'd', &time_scale },
  { "global_scale", 'd', &global_scale },
  { "steer_scale", 'd', &steer_scale },
  { "collision_scale", 'd', &collision_scale },
  { "match_scale", 'd', &match_scale },
  { "max_speed", 'd', &max_speed },
  { "num_agents", 'i', &num_agents },
  { "num_timesteps", 'i', &num_timesteps },
  { "min_pos", 'd', &min_pos },
  { "max_pos", 'd', &max_pos },
  { NULL, 0, NULL }
};
static const orc_agent agent_table[] = {
  { "Boid", Boid_info, &agents.agents_Boid, &agents.agents_Boid_dbuf, sizeof(Boid),
    offsetof(Boid, pos), 2, &grid_Boid },
};
static const orc_step step_table[] = {
  { "update_boid", 0, (void (*)(void *, void *)) update_boid, 0, radius_update_boid },
};
const orc_model_desc orc_model = {
  params, agent_table, 1, step_table, 1, model_env, model_init, agents_info, &agents,
};

#ifdef ORACLE_MAIN
#include "../orc_main.h"
ORC_DEFINE_MAIN("boids.out")
#endif
