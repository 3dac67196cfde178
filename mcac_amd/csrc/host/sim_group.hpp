// Lock-step driving of simulations whose boxes share a group (mcac_group): the
// ensemble runner picks up to `depth` steps ahead for every box (as the
// pipelined loop does), enqueues them together (one launch, one block per box)
// and reads every box's results, up to `launches` launches behind (with 2, it
// reads a launch's results while the next one runs).  Internal to the host
// library.
#pragma once
#include <cstdint>

#include "../../../include/mcac_gpu.h"
#include "../../../include/mcac_sim.h"

namespace mcac {

struct GroupMove {
    int64_t agg;
    double dir[3];
    double lpm;
    double maxradius;
};

// before mcac_sim_init; depth 1..8 steps per launch, launches 1..2 in flight
int sim_join_group(mcac_sim *sim, mcac_group *group, int depth, int launches = 1);
// The next step's move when one is due (not finished, fewer than max_steps
// done or in flight, fewer than depth in flight); it joins the flight queue.
bool sim_next_move(mcac_sim *sim, int64_t steps_done, int64_t max_steps, GroupMove *move);
mcac_box *sim_box(mcac_sim *sim);
// The ticket of the oldest move returned by sim_next_move and not yet ticketed.
void sim_issued(mcac_sim *sim, uint64_t ticket);
size_t sim_in_flight(mcac_sim *sim);
// Reads the issued step and runs the rest of the loop body; *stepped is false
// when the step did nothing and will be issued again.
int sim_consume(mcac_sim *sim, bool *stepped);
int sim_finalize(mcac_sim *sim);

}  // namespace mcac
