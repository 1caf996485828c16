/*
 * formats.cpp -- out/stats.csv (src/worker.cpp:63-181, src/timer.cpp:48-58).
 */
#include "formats.hpp"

namespace mcgpu {

void write_stats_csv(FILE *f, const std::vector<std::vector<TimerWindow>> &rank_windows) {
  /* the header is rank 0's (src/worker.cpp:75-83) */
  const bool timer_cols = !rank_windows.empty() && !rank_windows[0].empty();
  fprintf(f, "rank, ");
  if (timer_cols)
    fprintf(f, "%s, %s, %s, %s, %s, %s, ", "starttime", "endtime", "time_comp", "time_send",
            "time_recv", "time_idle");
  fprintf(f, "nb_cycles, \n");
  for (size_t r = 0; r < rank_windows.size(); ++r)
    for (const TimerWindow &w : rank_windows[r]) {
      fprintf(f, "%d, ", (int)r);
      fprintf(f, "%.18e, %.18e, %.18e, %.18e, %.18e, %.18e, ", w.starttime, w.endtime, w.time_comp,
              w.time_send, w.time_recv, w.time_idle);
      fprintf(f, "%d, \n", (int)w.nb_cycles);
    }
}

} // namespace mcgpu
