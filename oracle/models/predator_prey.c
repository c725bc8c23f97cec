/*
 * predator_prey.c -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * examples/predator_prey.abl written as the C backend would print it
 * (src/backend/CPrinter.cpp, GenericCPrinter.cpp) if it accepted run-time
 * add/remove, with the -b hip semantics of oracle/orc_dynamic.c for what the
 * C backend cannot express: removeCurrent(), add() inside steps, random()
 * inside steps (counter-based), and the sequential step gather_stats.
 * The C backend itself rejects this model (src/backend/CBackend.cpp:30-31),
 * so no reference run exists; the oracle pins the semantics the GPU path
 * must reproduce.
 */
#include ABL_RT_HEADER
#include "../orc_near.h"
#include "../orc_model.h"

#include <stdio.h>

typedef struct {
  float2 pos;
  float2 dir;
  float2 steer;
  int life;
} Predator;
static const type_info Predator_info[] = {
  { TYPE_FLOAT2, offsetof(Predator, pos), "pos", true },
  { TYPE_FLOAT2, offsetof(Predator, dir), "dir", false },
  { TYPE_FLOAT2, offsetof(Predator, steer), "steer", false },
  { TYPE_INT, offsetof(Predator, life), "life", false },
  { TYPE_END, sizeof(Predator), NULL }
};
typedef struct {
  float2 pos;
  float2 dir;
  float2 steer;
  int life;
} Prey;
static const type_info Prey_info[] = {
  { TYPE_FLOAT2, offsetof(Prey, pos), "pos", true },
  { TYPE_FLOAT2, offsetof(Prey, dir), "dir", false },
  { TYPE_FLOAT2, offsetof(Prey, steer), "steer", false },
  { TYPE_INT, offsetof(Prey, life), "life", false },
  { TYPE_END, sizeof(Prey), NULL }
};
typedef struct {
  float2 pos;
  int dead_cycles;
  bool avail;
} Grass;
static const type_info Grass_info[] = {
  { TYPE_FLOAT2, offsetof(Grass, pos), "pos", true },
  { TYPE_INT, offsetof(Grass, dead_cycles), "dead_cycles", false },
  { TYPE_BOOL, offsetof(Grass, avail), "avail", false },
  { TYPE_END, sizeof(Grass), NULL }
};
struct agent_struct {
  dyn_array agents_Predator;
  dyn_array agents_Predator_dbuf;
  dyn_array agents_Prey;
  dyn_array agents_Prey_dbuf;
  dyn_array agents_Grass;
  dyn_array agents_Grass_dbuf;
};
struct agent_struct agents;
static const agent_info agents_info[] = {
  { Predator_info, offsetof(struct agent_struct, agents_Predator), "Predator" },
  { Prey_info, offsetof(struct agent_struct, agents_Prey), "Prey" },
  { Grass_info, offsetof(struct agent_struct, agents_Grass), "Grass" },
  { NULL, 0, NULL }
};

enum { T_PREDATOR, T_PREY, T_GRASS };

abl_float PI = 3.14159;
abl_float SPACE_MULT = 1.0;
abl_float REPRODUCE_PREY_PROB = 0.005;
abl_float REPRODUCE_PREDATOR_PROB = 0.001;
int GAIN_FROM_FOOD_PREDATOR = 35;
int GAIN_FROM_FOOD_PREY = 60;
int GRASS_REGROW_CYCLES = 60;
abl_float PRED_PREY_INTERACTION_RADIUS = 0.1;
abl_float PREY_GROUP_COHESION_RADIUS = 0.12;
abl_float SAME_SPECIES_AVOIDANCE_RADIUS = 0.035;
abl_float GRASS_EAT_DISTANCE = 0.02;
abl_float PRED_KILL_DISTANCE = 0.02;
abl_float DELTA_TIME = 0.0015;
abl_float PRED_SPEED_ADVANTAGE = 2.0;
abl_float prey_fraction = 0.35;
abl_float predator_fraction = 0.025;
abl_float agent_density = 1600.0;
int num_agents = 3200;
int num_timesteps = 100;
abl_float env_size = 1.41421;

static orc_grid grid_Predator, grid_Prey, grid_Grass;

float2 random_1(float2 min, float2 max) {
  return float2_create(random_float(min.x, max.x), random_float(min.y, max.y));
}
abl_float random_3(abl_float max) { return random_float(0, max); }
float2 random_4(float2 max) { return random_1(float2_fill(0), max); }
int randomInt_1(int max) { return random_int(0, max); }
float2 clamp_1(float2 pos, float2 min, float2 max) {
  return float2_create(((pos.x < min.x) ? min.x : ((pos.x > max.x) ? max.x : pos.x)),
                       ((pos.y < min.y) ? min.y : ((pos.y > max.y) ? max.y : pos.y)));
}
float2 clamp_4(float2 pos, float2 max) { return clamp_1(pos, float2_fill(0), max); }

/* random() inside a step: the counter-based generator */
static abl_float step_random_3(abl_float max) { return orc_step_random(0, max); }

float2 randomDirection(void) {
  abl_float phi = random_3((2.0 * PI));
  return float2_create(sin(phi), cos(phi));
}

float2 boundPosition(float2 pos) {
  return clamp_4(pos, float2_create(env_size, env_size));
}

void prey_avoid_pred(Prey* in, Prey* out) {
  float2 avoid_velocity = float2_fill(0.0);
  ORC_NEAR_OPEN(Predator, pred, agents.agents_Predator, grid_Predator, in->pos)
    if (dist_float2(pred->pos, in->pos) > PRED_PREY_INTERACTION_RADIUS) continue;
    ORC_NEAR_HIT(pred, agents.agents_Predator, Predator)
    {
      abl_float separation = length_float2(float2_sub(in->pos, pred->pos));
      if ((separation < PRED_PREY_INTERACTION_RADIUS)) {
        if ((separation > 0.0)) {
          avoid_velocity = float2_add(avoid_velocity, float2_mul_scalar(float2_sub(in->pos, pred->pos), (PRED_PREY_INTERACTION_RADIUS / separation)));
        }
      }
    }
  ORC_NEAR_CLOSE
  out->steer = avoid_velocity;
}

void prey_flock(Prey* in, Prey* out) {
  float2 group_center = float2_fill(0.0);
  float2 group_velocity = float2_fill(0.0);
  float2 avoid_velocity = float2_fill(0.0);
  int group_centre_count = 0;
  ORC_NEAR_OPEN(Prey, prey, agents.agents_Prey, grid_Prey, in->pos)
    if (dist_float2(prey->pos, in->pos) > PREY_GROUP_COHESION_RADIUS) continue;
    ORC_NEAR_HIT(prey, agents.agents_Prey, Prey)
    {
      abl_float separation = length_float2(float2_sub(in->pos, prey->pos));
      group_center = float2_add(group_center, prey->pos);
      group_centre_count += 1;
      if (((separation < SAME_SPECIES_AVOIDANCE_RADIUS) && float2_not_equals(prey->pos, in->pos))) {
        if ((separation > 0.0)) {
          avoid_velocity = float2_add(avoid_velocity, float2_mul_scalar(float2_sub(in->pos, prey->pos), (SAME_SPECIES_AVOIDANCE_RADIUS / separation)));
        }
      }
    }
  ORC_NEAR_CLOSE
  if ((group_centre_count > 0)) {
    group_center = float2_div_scalar(group_center, group_centre_count);
    group_velocity = float2_sub(group_center, in->pos);
  }
  out->steer = float2_add(float2_add(in->steer, group_velocity), avoid_velocity);
}

void prey_move(Prey* in, Prey* out) {
  float2 agent_position = in->pos;
  float2 agent_velocity = in->dir;
  float2 agent_steer = in->steer;
  agent_velocity = float2_add(agent_velocity, agent_steer);
  abl_float current_speed = length_float2(agent_velocity);
  if ((current_speed > 1.0)) agent_velocity = normalize_float2(agent_velocity);
  agent_position = float2_add(agent_position, float2_mul_scalar(float2_mul_scalar(agent_velocity, DELTA_TIME), SPACE_MULT));
  agent_position = boundPosition(agent_position);
  out->pos = agent_position;
  out->dir = agent_velocity;
  out->life = (in->life - 1);
}

void prey_eaten(Prey* in, Prey* out) {
  bool eaten = false;
  abl_float closest_pred = PRED_KILL_DISTANCE;
  ORC_NEAR_OPEN(Predator, pred, agents.agents_Predator, grid_Predator, in->pos)
    if (dist_float2(pred->pos, in->pos) > PRED_KILL_DISTANCE) continue;
    ORC_NEAR_HIT(pred, agents.agents_Predator, Predator)
    {
      abl_float distance = length_float2(float2_sub(pred->pos, in->pos));
      if ((distance < closest_pred)) {
        closest_pred = distance;
        eaten = true;
      }
    }
  ORC_NEAR_CLOSE
  if (eaten) orc_remove_current();
}

void prey_eat_or_starve(Prey* in, Prey* out) {
  int life = in->life;
  ORC_NEAR_OPEN(Grass, grass, agents.agents_Grass, grid_Grass, in->pos)
    if (dist_float2(grass->pos, in->pos) > GRASS_EAT_DISTANCE) continue;
    ORC_NEAR_HIT(grass, agents.agents_Grass, Grass)
    {
      if (grass->avail) life += GAIN_FROM_FOOD_PREY;
    }
  ORC_NEAR_CLOSE
  out->life = life;
  if ((life < 1)) orc_remove_current();
}

void prey_reproduction(Prey* in, Prey* out) {
  if ((step_random_3(1.0) < REPRODUCE_PREY_PROB)) {
    Prey child = (Prey) {
      .pos = in->pos,
      .dir = float2_mul_scalar(in->dir, -1.0),
      .steer = float2_mul_scalar(in->steer, -1.0),
      .life = (in->life / 2),
    };
    orc_add_agent(T_PREY, &child);
    out->life = (in->life / 2);
  }
}

void pred_follow_prey(Predator* in, Predator* out) {
  float2 agent_position = in->pos;
  float2 agent_steer = float2_create(0, 0);
  float2 closest_prey_position = float2_create(0, 0);
  abl_float closest_prey_distance = PRED_PREY_INTERACTION_RADIUS;
  bool can_see_prey = false;
  ORC_NEAR_OPEN(Prey, prey, agents.agents_Prey, grid_Prey, in->pos)
    if (dist_float2(prey->pos, in->pos) > PRED_PREY_INTERACTION_RADIUS) continue;
    ORC_NEAR_HIT(prey, agents.agents_Prey, Prey)
    {
      abl_float separation = length_float2(float2_sub(in->pos, prey->pos));
      if ((separation < closest_prey_distance)) {
        closest_prey_position = prey->pos;
        closest_prey_distance = separation;
        can_see_prey = true;
      }
    }
  ORC_NEAR_CLOSE
  if (can_see_prey) {
    agent_steer = float2_sub(closest_prey_position, agent_position);
  }
  out->steer = agent_steer;
}

void pred_avoid(Predator* in, Predator* out) {
  float2 avoid_velocity = float2_create(0, 0);
  ORC_NEAR_OPEN(Predator, pred, agents.agents_Predator, grid_Predator, in->pos)
    if (dist_float2(pred->pos, in->pos) > SAME_SPECIES_AVOIDANCE_RADIUS) continue;
    ORC_NEAR_HIT(pred, agents.agents_Predator, Predator)
    {
      abl_float separation = length_float2(float2_sub(in->pos, pred->pos));
      if (((separation < SAME_SPECIES_AVOIDANCE_RADIUS) && float2_not_equals(pred->pos, in->pos))) {
        if ((separation > 0.0)) {
          avoid_velocity = float2_add(avoid_velocity, float2_mul_scalar(float2_sub(in->pos, pred->pos), (SAME_SPECIES_AVOIDANCE_RADIUS / separation)));
        }
      }
    }
  ORC_NEAR_CLOSE
  out->steer = float2_add(in->steer, avoid_velocity);
}

void pred_move(Predator* in, Predator* out) {
  float2 agent_position = in->pos;
  float2 agent_velocity = in->dir;
  float2 agent_steer = in->steer;
  agent_velocity = float2_add(agent_velocity, agent_steer);
  abl_float current_speed = length_float2(agent_velocity);
  if ((current_speed > 1.0)) {
    agent_velocity = normalize_float2(agent_velocity);
  }
  agent_position = float2_add(agent_position, float2_mul_scalar(float2_mul_scalar(float2_mul_scalar(agent_velocity, DELTA_TIME), PRED_SPEED_ADVANTAGE), SPACE_MULT));
  agent_position = boundPosition(agent_position);
  out->pos = agent_position;
  out->dir = agent_velocity;
  out->life = (in->life - 1);
}

void pred_eat_or_starve(Predator* in, Predator* out) {
  int pred_life = in->life;
  ORC_NEAR_OPEN(Prey, prey, agents.agents_Prey, grid_Prey, in->pos)
    if (dist_float2(prey->pos, in->pos) > PRED_KILL_DISTANCE) continue;
    ORC_NEAR_HIT(prey, agents.agents_Prey, Prey)
    {
      pred_life += GAIN_FROM_FOOD_PREDATOR;
    }
  ORC_NEAR_CLOSE
  out->life = pred_life;
  if ((pred_life < 1)) orc_remove_current();
}

void pred_reproduction(Predator* in, Predator* out) {
  if ((step_random_3(1.0) < REPRODUCE_PREDATOR_PROB)) {
    Predator child = (Predator) {
      .pos = in->pos,
      .dir = float2_mul_scalar(in->dir, -1.0),
      .steer = float2_mul_scalar(in->steer, -1.0),
      .life = (in->life / 2),
    };
    orc_add_agent(T_PREDATOR, &child);
    out->life = (in->life / 2);
  }
}

void grass_eaten(Grass* in, Grass* out) {
  abl_float closest_prey = GRASS_EAT_DISTANCE;
  bool eaten = false;
  ORC_NEAR_OPEN(Prey, prey, agents.agents_Prey, grid_Prey, in->pos)
    if (dist_float2(prey->pos, in->pos) > GRASS_EAT_DISTANCE) continue;
    ORC_NEAR_HIT(prey, agents.agents_Prey, Prey)
    {
      abl_float distance = length_float2(float2_sub(in->pos, prey->pos));
      if (((distance < closest_prey) && in->avail)) {
        closest_prey = distance;
        eaten = true;
      }
    }
  ORC_NEAR_CLOSE
  if (eaten) {
    out->dead_cycles = 0;
    out->avail = false;
  }
}

void grass_growth(Grass* in, Grass* out) {
  if ((in->dead_cycles == GRASS_REGROW_CYCLES)) {
    int cycle_start = 0;
    out->dead_cycles = cycle_start;
    out->avail = true;
  }
  if ((!in->avail)) out->dead_cycles = (in->dead_cycles + 1);
}

/* sequential step gather_stats(): count, sum, getLastExecTime, log_csv
 * (log_csv prints "%d"/"%f" joined by ',' as FlameGPUPrinter.cpp:226-245) */
void gather_stats(void) {
  int num_prey = orc_count_type(T_PREY);
  int num_predator = orc_count_type(T_PREDATOR);
  int num_avail_grass = (int) orc_sum_member_int(T_GRASS, offsetof(Grass, avail), 1);
  abl_float exec_time = orc_last_exec_time();
  char row[128];
  snprintf(row, sizeof row, "%d,%d,%d,%f\n", num_prey, num_predator, num_avail_grass,
           (double) exec_time);
  orc_log_csv_row(row);
}

static void model_init(void) {
  int num_predators = ((int) (predator_fraction * ((abl_float) num_agents)));
  int num_prey = ((int) (prey_fraction * ((abl_float) num_agents)));
  int num_grass = ((num_agents - num_predators) - num_prey);
  for (int i = 0, _var1 = num_grass; i < _var1; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Grass, Grass) = (Grass) {
      .pos = random_4(float2_fill(env_size)),
      .dead_cycles = 0,
      .avail = true,
    };
  }
  for (int i = 0, _var2 = num_prey; i < _var2; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Prey, Prey) = (Prey) {
      .pos = random_4(float2_fill(env_size)),
      .dir = randomDirection(),
      .steer = float2_fill(0),
      .life = (GAIN_FROM_FOOD_PREY + randomInt_1(10)),
    };
  }
  for (int i = 0, _var3 = num_predators; i < _var3; ++i) {
    *DYN_ARRAY_PLACE(&agents.agents_Predator, Predator) = (Predator) {
      .pos = random_4(float2_fill(env_size)),
      .dir = randomDirection(),
      .steer = float2_fill(0),
      .life = (GAIN_FROM_FOOD_PREDATOR + randomInt_1(10)),
    };
  }
}

static double r_interact(void) { return (double) PRED_PREY_INTERACTION_RADIUS; }
static double r_cohesion(void) { return (double) PREY_GROUP_COHESION_RADIUS; }
static double r_avoid(void) { return (double) SAME_SPECIES_AVOIDANCE_RADIUS; }
static double r_grass(void) { return (double) GRASS_EAT_DISTANCE; }
static double r_kill(void) { return (double) PRED_KILL_DISTANCE; }

static void model_env(double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = (double) env_size;
  hi[2] = 0.0;
  /* granularity: the largest constant near radius (AnalysisVisitor.cpp:1378-1399) */
  double m = r_interact();
  if (r_cohesion() > m) m = r_cohesion();
  if (r_avoid() > m) m = r_avoid();
  if (r_grass() > m) m = r_grass();
  if (r_kill() > m) m = r_kill();
  *g = m;
}

static const orc_param params[] = {
  { "SPACE_MULT", 'd', &SPACE_MULT },
  { "REPRODUCE_PREY_PROB", 'd', &REPRODUCE_PREY_PROB },
  { "REPRODUCE_PREDATOR_PROB", 'd', &REPRODUCE_PREDATOR_PROB },
  { "GAIN_FROM_FOOD_PREDATOR", 'i', &GAIN_FROM_FOOD_PREDATOR },
  { "GAIN_FROM_FOOD_PREY", 'i', &GAIN_FROM_FOOD_PREY },
  { "GRASS_REGROW_CYCLES", 'i', &GRASS_REGROW_CYCLES },
  { "PRED_PREY_INTERACTION_RADIUS", 'd', &PRED_PREY_INTERACTION_RADIUS },
  { "PREY_GROUP_COHESION_RADIUS", 'd', &PREY_GROUP_COHESION_RADIUS },
  { "SAME_SPECIES_AVOIDANCE_RADIUS", 'd', &SAME_SPECIES_AVOIDANCE_RADIUS },
  { "GRASS_EAT_DISTANCE", 'd', &GRASS_EAT_DISTANCE },
  { "PRED_KILL_DISTANCE", 'd', &PRED_KILL_DISTANCE },
  { "DELTA_TIME", 'd', &DELTA_TIME },
  { "PRED_SPEED_ADVANTAGE", 'd', &PRED_SPEED_ADVANTAGE },
  { "prey_fraction", 'd', &prey_fraction },
  { "predator_fraction", 'd', &predator_fraction },
  { "agent_density", 'd', &agent_density },
  { "num_agents", 'i', &num_agents },
  { "num_timesteps", 'i', &num_timesteps },
  { "env_size", 'd', &env_size },
  { NULL, 0, NULL }
};
static const orc_agent agent_table[] = {
  { "Predator", Predator_info, &agents.agents_Predator, &agents.agents_Predator_dbuf,
    sizeof(Predator), offsetof(Predator, pos), 2, &grid_Predator },
  { "Prey", Prey_info, &agents.agents_Prey, &agents.agents_Prey_dbuf, sizeof(Prey),
    offsetof(Prey, pos), 2, &grid_Prey },
  { "Grass", Grass_info, &agents.agents_Grass, &agents.agents_Grass_dbuf, sizeof(Grass),
    offsetof(Grass, pos), 2, &grid_Grass },
};
#define STEP(fn, agent, near, radius, removes, adds) \
  { #fn, agent, (void (*)(void *, void *)) fn, near, radius, removes, adds }
static const orc_step step_table[] = {
  STEP(pred_follow_prey, T_PREDATOR, T_PREY, r_interact, 0, 0),
  STEP(prey_avoid_pred, T_PREY, T_PREDATOR, r_interact, 0, 0),
  STEP(prey_flock, T_PREY, T_PREY, r_cohesion, 0, 0),
  STEP(pred_avoid, T_PREDATOR, T_PREDATOR, r_avoid, 0, 0),
  STEP(prey_move, T_PREY, -1, NULL, 0, 0),
  STEP(pred_move, T_PREDATOR, -1, NULL, 0, 0),
  STEP(prey_eat_or_starve, T_PREY, T_GRASS, r_grass, 1, 0),
  STEP(pred_eat_or_starve, T_PREDATOR, T_PREY, r_kill, 1, 0),
  STEP(grass_eaten, T_GRASS, T_PREY, r_grass, 0, 0),
  STEP(prey_eaten, T_PREY, T_PREDATOR, r_kill, 1, 0),
  STEP(pred_reproduction, T_PREDATOR, -1, NULL, 0, 1 + T_PREDATOR),
  STEP(prey_reproduction, T_PREY, -1, NULL, 0, 1 + T_PREY),
  STEP(grass_growth, T_GRASS, -1, NULL, 0, 0),
};
const orc_model_desc orc_model = {
  params, agent_table, 3, step_table, 13, model_env, model_init, agents_info, &agents,
  1, gather_stats,
};

#ifdef ORACLE_MAIN
#include "../orc_main.h"
ORC_DEFINE_MAIN("agents.json")
#endif
