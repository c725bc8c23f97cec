/*
 * game_of_life.c -- TEST INFRASTRUCTURE ONLY (oracle).
 *
 * Hand restatement of what the reference C backend emits for
 * examples/game_of_life.abl: a grid of Cell agents at (x+0.5, y+0.5), near
 * radius 1.5 (the 8-neighbourhood plus the cell itself), B3/S23 rule, the
 * random initialisation `random(0, 1) < alive_fraction` drawn in x-major,
 * y-minor order.  getColor() is not emitted for the C backend
 * (src/AnalysisVisitor.cpp:215-222).
 */
#include ABL_RT_HEADER
#include "../orc_near.h"
#include "../orc_model.h"

typedef struct {
  float2 pos;
  bool alive;
} Cell;
static const type_info Cell_info[] = {
  { TYPE_FLOAT2, offsetof(Cell, pos), "pos", true },
  { TYPE_BOOL, offsetof(Cell, alive), "alive", false },
  { TYPE_END, sizeof(Cell), NULL }
};
struct agent_struct {
  dyn_array agents_Cell;
  dyn_array agents_Cell_dbuf;
};
struct agent_struct agents;
static const agent_info agents_info[] = {
  { Cell_info, offsetof(struct agent_struct, agents_Cell), "Cell" },
  { NULL, 0, NULL }
};

abl_float PI = 3.14159;
abl_float radius = 1.5;
bool debug = false;
abl_float alive_fraction = 0.1;
int num_agents = 400;
int num_timesteps = 100;
int size = 20;

static orc_grid grid_Cell;

void step_move(Cell* in, Cell* out) {
  int living_neighbors = 0;
  ORC_NEAR_OPEN(Cell, nx, agents.agents_Cell, grid_Cell, in->pos)
    if (dist_float2(nx->pos, in->pos) > radius) continue;
    ORC_NEAR_HIT(nx, agents.agents_Cell, Cell)
    {
      living_neighbors += (nx->alive ? 1 : 0);
    }
  ORC_NEAR_CLOSE
  if (in->alive) {
    living_neighbors -= 1;
    out->alive = ((living_neighbors >= 2) && (living_neighbors <= 3));
  } else {
    out->alive = (living_neighbors == 3);
  }
  out->pos = in->pos;
}

bool getInit(int x, int y) {
  if (debug) {
    return ((((x == 10) || (x == 11)) || (x == 12)) && (y == 10));
  } else {
    return (random_float(0, 1) < alive_fraction);
  }
}

static void model_init(void) {
  for (int x = 0, _var1 = size; x < _var1; ++x) {
    for (int y = 0, _var2 = size; y < _var2; ++y) {
      *DYN_ARRAY_PLACE(&agents.agents_Cell, Cell) = (Cell) {
        .pos = float2_create((((abl_float) x) + 0.5), (((abl_float) y) + 0.5)),
        .alive = getInit(x, y),
      };
    }
  }
}

static double radius_step_move(void) { return (double) radius; }

static void model_env(double lo[3], double hi[3], double *g) {
  lo[0] = lo[1] = lo[2] = 0.0;
  hi[0] = hi[1] = (double) size;
  hi[2] = 0.0;
  *g = (double) radius;
}

static const orc_param params[] = {
  { "debug", 'b', &debug },
  { "alive_fraction", 'd', &alive_fraction },
  { "num_agents", 'i', &num_agents },
  { "num_timesteps", 'i', &num_timesteps },
  { "size", 'i', &size },
  { NULL, 0, NULL }
};
static const orc_agent agent_table[] = {
  { "Cell", Cell_info, &agents.agents_Cell, &agents.agents_Cell_dbuf, sizeof(Cell),
    offsetof(Cell, pos), 2, &grid_Cell },
};
static const orc_step step_table[] = {
  { "step_move", 0, (void (*)(void *, void *)) step_move, 0, radius_step_move },
};
const orc_model_desc orc_model = {
  params, agent_table, 1, step_table, 1, model_env, model_init, agents_info, &agents,
};

#ifdef ORACLE_MAIN
#include "../orc_main.h"
ORC_DEFINE_MAIN("result.json")
#endif
