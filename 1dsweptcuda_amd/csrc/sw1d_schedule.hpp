// The reference programs' time bookkeeping, replayed on the host.
//
// The reference advances in whole cycles and prints a .dat row whenever the
// accumulated time passes the next output time, so the number of steps taken
// and the printed times are quantised by the cycle length.  These loops repeat
// its double-precision arithmetic operation for operation (repeated addition,
// the 0.25*dt guards, the comparison operators) and record, for every row the
// program writes, the number of steps computed so far and the time printed.
#pragma once

#include <cstdint>
#include <vector>

namespace sw1d {

enum SchedKind : int {
    SCHED_HEAT_CLASSIC = 0,  // Heat:483-510
    SCHED_HEAT_SWEPT = 1,    // Heat:541-682 (GPU swept and hybrid)
    SCHED_KS_CLASSIC = 2,    // KS:609-631
    SCHED_KS_SWEPT = 3,      // KS:676, 804-851
    SCHED_KS_REGISTER = 4,   // KS:753-802 (node fixed at 32 points)
    SCHED_EULER_CLASSIC = 5, // Euler:855-888
    SCHED_EULER_SWEPT = 6    // Euler:924-1098 (GPU swept and hybrid)
};

struct SchedRow {
    int64_t steps;  // steps computed when the row is written
    double time;    // time value printed in the row
};

// tf_arg: argv[4]; the programs end at t_end = tf_arg - 0.25*dt (Heat:715,
// KS:883, Euler:1140).  dt_sched: the dt the wrapper receives; KS passes the
// REAL-rounded dsc.dt (KS:975, 979), the others pass the argv value.
// The last row is the final state.
inline bool run_schedule(int kind, double dt, double tf_arg, double freq, int tpb,
                         double dt_sched, std::vector<SchedRow>& rows)
{
    rows.clear();
    const double t_end = tf_arg - 0.25 * dt;
    const double d = dt_sched;
    double twrite = freq - 0.25 * d;
    double t_eq = 0.0;
    int64_t n = 0;
    switch (kind) {
    case SCHED_HEAT_CLASSIC: {
        const double full = d + d;
        n = 2;
        t_eq = full;
        while (t_eq < t_end) {
            n += 2;
            t_eq += full;
            if (t_eq > twrite) {
                rows.push_back({n, t_eq});
                t_eq += full;  // advances the clock without computing (Heat:506)
                twrite += freq;
            }
        }
        break;
    }
    case SCHED_KS_CLASSIC: {
        while (t_eq <= t_end) {
            n += 1;
            t_eq += d;
            if (t_eq > twrite) {
                rows.push_back({n, t_eq});
                twrite += freq;
            }
        }
        break;
    }
    case SCHED_EULER_CLASSIC: {
        while (t_eq < t_end) {
            n += 1;
            t_eq += d;
            if (t_eq > twrite) {
                rows.push_back({n, t_eq});
                twrite += freq;
            }
        }
        break;
    }
    case SCHED_HEAT_SWEPT:
    case SCHED_KS_SWEPT:
    case SCHED_KS_REGISTER:
    case SCHED_EULER_SWEPT: {
        double full;
        int64_t per;
        if (kind == SCHED_HEAT_SWEPT) {
            full = d * (double)tpb;
            per = tpb;
        } else if (kind == SCHED_KS_REGISTER) {
            full = 0.25 * d * 32.0;
            per = 8;
        } else {
            full = 0.25 * d * (double)tpb;
            per = tpb / 4;
        }
        if (per <= 0) return false;
        // up-triangle + first split diamond; realised by the final down-triangle.
        n = per;
        t_eq = full;
        while (t_eq < t_end) {
            n += per;
            t_eq += full;
            if (t_eq > twrite) {
                rows.push_back({n, t_eq});
                // the output re-seeds the pipeline with up + split (one cycle)
                n += per;
                t_eq += full;
                twrite += freq;
            }
        }
        break;
    }
    default:
        return false;
    }
    rows.push_back({n, t_eq});
    return true;
}

}  // namespace sw1d
