/*
 * ref_probe.c -- TEST INFRASTRUCTURE ONLY.
 *
 * Prints exact outputs of the C runtime so the oracle can be pinned against
 * the reference: built against the reference's asset/c/libabl.c (oracle/_ref)
 * and against the oracle's own abl_rt.c (oracle/_build); the tests require the
 * two to agree byte for byte and commit the reference's output as fixtures.
 *
 *   ref_probe float <count> <lo> <hi>      random_float draws, hex (%a)
 *   ref_probe int <count> <lo> <hi>        random_int draws
 *   ref_probe mix <count>                  interleaved float/int draws
 *   ref_probe dist2 <file>                 dist_float2 for "ax ay bx by" lines (hex in/out)
 *   ref_probe dist3 <file>                 dist_float3 for 6-value lines
 *   ref_probe save <path>                  save() of a fixed two-type record set
 *   ref_probe save_flame <json> <xml>      the same records as JSON, then as FLAME XML
 *   ref_probe save_flamegpu <json> <xml>   the same, FlameGPU XML
 */
#ifdef ABL_RT_HEADER
#include ABL_RT_HEADER
#else
#include "libabl.h"
#endif
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  float2 pos;
  bool alive;
  int life;
  abl_float energy;
} ProbeA;
static const type_info ProbeA_info[] = {
  { TYPE_FLOAT2, offsetof(ProbeA, pos), "pos", true },
  { TYPE_BOOL, offsetof(ProbeA, alive), "alive", false },
  { TYPE_INT, offsetof(ProbeA, life), "life", false },
  { TYPE_FLOAT, offsetof(ProbeA, energy), "energy", false },
  { TYPE_END, sizeof(ProbeA), NULL }
};
typedef struct {
  float3 pos;
  float3 vel;
} ProbeB;
static const type_info ProbeB_info[] = {
  { TYPE_FLOAT3, offsetof(ProbeB, pos), "pos", true },
  { TYPE_FLOAT3, offsetof(ProbeB, vel), "vel", false },
  { TYPE_END, sizeof(ProbeB), NULL }
};
struct probe_agents {
  dyn_array agents_ProbeA;
  dyn_array agents_ProbeA_dbuf;
  dyn_array agents_ProbeB;
  dyn_array agents_ProbeB_dbuf;
};
static struct probe_agents probe;
static const agent_info probe_info[] = {
  { ProbeA_info, offsetof(struct probe_agents, agents_ProbeA), "ProbeA" },
  { ProbeB_info, offsetof(struct probe_agents, agents_ProbeB), "ProbeB" },
  { NULL, 0, NULL }
};

static void probe_records(void) {
  for (int i = 0; i < 5; i++) {
    ProbeA *a = DYN_ARRAY_PLACE(&probe.agents_ProbeA, ProbeA);
    a->pos = float2_create(random_float(-1000, 1000), random_float(0, 1));
    a->alive = (i % 2) == 0;
    a->life = random_int(-50, 50);
    a->energy = random_float(-1e-7, 1e-7) + (abl_float) i * 1234567.125;
  }
  for (int i = 0; i < 3; i++) {
    ProbeB *b = DYN_ARRAY_PLACE(&probe.agents_ProbeB, ProbeB);
    b->pos = float3_create(random_float(0, 1), random_float(0, 10), random_float(0, 100));
    b->vel = float3_create(-0.0000005, 0.0000005, -2.5);
  }
}

static int probe_save(const char *path, save_type type) {
  probe_records();
  save(&probe, probe_info, path, type);
  return 0;
}

/* the same records written as JSON and then as the given XML form */
static int probe_save_both(const char *json_path, const char *xml_path, save_type type) {
  probe_records();
  save(&probe, probe_info, json_path, SAVE_JSON);
  save(&probe, probe_info, xml_path, type);
  return 0;
}

static int probe_dist(const char *path, int dim) {
  FILE *f = fopen(path, "r");
  if (!f) return 1;
  double v[6];
  for (;;) {
    int got = 0;
    for (int k = 0; k < 2 * dim; k++) got += fscanf(f, "%la", &v[k]);
    if (got != 2 * dim) break;
    abl_float d;
    if (dim == 2) {
      d = dist_float2(float2_create(v[0], v[1]), float2_create(v[2], v[3]));
    } else {
      d = dist_float3(float3_create(v[0], v[1], v[2]), float3_create(v[3], v[4], v[5]));
    }
    printf("%a\n", (double) d);
  }
  fclose(f);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  const char *mode = argv[1];
  if (strcmp(mode, "float") == 0 && argc == 5) {
    int n = atoi(argv[2]);
    double lo = atof(argv[3]), hi = atof(argv[4]);
    for (int i = 0; i < n; i++) printf("%a\n", (double) random_float(lo, hi));
    return 0;
  }
  if (strcmp(mode, "int") == 0 && argc == 5) {
    int n = atoi(argv[2]);
    int lo = atoi(argv[3]), hi = atoi(argv[4]);
    for (int i = 0; i < n; i++) printf("%d\n", random_int(lo, hi));
    return 0;
  }
  if (strcmp(mode, "mix") == 0) {
    int n = atoi(argv[2]);
    for (int i = 0; i < n; i++) {
      printf("%a %d %d\n", (double) random_float(-3, 7), random_int(0, 9), random_int(5, 12));
    }
    return 0;
  }
  if (strcmp(mode, "dist2") == 0) return probe_dist(argv[2], 2);
  if (strcmp(mode, "dist3") == 0) return probe_dist(argv[2], 3);
  if (strcmp(mode, "save") == 0) return probe_save(argv[2], SAVE_JSON);
  if (strcmp(mode, "save_flame") == 0 && argc == 4) return probe_save_both(argv[2], argv[3], SAVE_FLAME_XML);
  if (strcmp(mode, "save_flamegpu") == 0 && argc == 4)
    return probe_save_both(argv[2], argv[3], SAVE_FLAMEGPU_XML);
  return 2;
}
