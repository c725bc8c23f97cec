/*
 * orc_main.h -- TEST INFRASTRUCTURE ONLY.
 *
 * main() of a restated generated program: initialisation, the simulate loop,
 * then save() (src/backend/CPrinter.cpp:267-272, 74-77).  The simulate loop
 * alone is timed with clock_gettime and printed as
 *   "simulate_seconds=<s> agent_steps=<n>"
 * so the CPU baseline excludes init and the JSON write.  Parameters arrive as
 * name=value arguments (the effect of -P at code-generation time); a
 * "timesteps=<n>" argument bounds the number of timed steps for sampling.
 */
#ifndef ORACLE_ORC_MAIN_H
#define ORACLE_ORC_MAIN_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

int orc_set_param(const char *name, double value);
double orc_get_param(const char *name);
void orc_simulate(int timesteps);
void orc_init(void);
void orc_save(const char *path);
size_t orc_agent_count(int t);
int orc_num_agent_types(void);

void orc_sample_step(int s, size_t m);

/* "sample=<m>": time step function 0 for the first m agents only (each
 * still scanning the whole population) and print
 *   "sample_seconds=<s> sample_agents=<m>". */
#define ORC_DEFINE_MAIN(default_out)                                           \
  int main(int argc, char **argv) {                                            \
    const char *out = default_out;                                             \
    int steps = -1;                                                            \
    long sample = -1;                                                          \
    for (int i = 1; i < argc; i++) {                                           \
      char *eq = strchr(argv[i], '=');                                         \
      if (!eq) continue;                                                       \
      *eq = 0;                                                                 \
      if (strcmp(argv[i], "out") == 0) { out = eq + 1; continue; }             \
      if (strcmp(argv[i], "timesteps") == 0) { steps = atoi(eq + 1); continue; } \
      if (strcmp(argv[i], "sample") == 0) { sample = atol(eq + 1); continue; } \
      if (orc_set_param(argv[i], atof(eq + 1)) != 0) {                         \
        fprintf(stderr, "unknown parameter %s\n", argv[i]);                    \
        return 1;                                                              \
      }                                                                        \
    }                                                                          \
    if (steps < 0) steps = (int) orc_get_param("num_timesteps");              \
    orc_init();                                                                \
    if (sample > 0) {                                                          \
      struct timespec s0, s1;                                                  \
      clock_gettime(CLOCK_MONOTONIC, &s0);                                     \
      orc_sample_step(0, (size_t) sample);                                     \
      clock_gettime(CLOCK_MONOTONIC, &s1);                                     \
      printf("sample_seconds=%.9f sample_agents=%ld population=%zu\n",         \
             (s1.tv_sec - s0.tv_sec) + 1e-9 * (s1.tv_nsec - s0.tv_nsec), sample, \
             orc_agent_count(0));                                              \
      return 0;                                                                \
    }                                                                          \
    size_t alive = 0;                                                          \
    for (int t = 0; t < orc_num_agent_types(); t++) alive += orc_agent_count(t); \
    struct timespec t0, t1;                                                    \
    clock_gettime(CLOCK_MONOTONIC, &t0);                                       \
    orc_simulate(steps);                                                       \
    clock_gettime(CLOCK_MONOTONIC, &t1);                                       \
    double secs = (t1.tv_sec - t0.tv_sec) + 1e-9 * (t1.tv_nsec - t0.tv_nsec);  \
    printf("simulate_seconds=%.9f agent_steps=%zu\n", secs, alive * (size_t) steps); \
    if (strcmp(out, "-") != 0) orc_save(out);                                  \
    return 0;                                                                  \
  }

#endif
