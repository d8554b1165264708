// TEST INFRASTRUCTURE: host emulation of the mixed-batch launcher (see fx_launch.hpp for the
// emulation rules).  The grid is 3 workgroups walking the task list round-robin; each task
// runs its three phases over all threads in barrier order, LDS re-poisoned per task.
#pragma once

namespace fx {

template <typename T, int PRESET>
hipError_t launch_multi(const void* args, const int2* tasks, int ntasks, const void* bank,
                        unsigned* work, hipStream_t)
{
    using M = MultiCfg<T, PRESET>;
    const MultiArgs<T>& a = *(const MultiArgs<T>*)args;
    const Bank<T>* bk = (const Bank<T>*)bank;
    if (ntasks <= 0) return hipSuccess;
    EmuLds<T> lds(M::LDS);
    const long grid = ntasks < kEmuPersistentGrid ? ntasks : kEmuPersistentGrid;
    emu_persistent(
        ntasks, grid, ntasks >= 2 * grid ? work : nullptr, [](long, long) {},
        [&](long, long t, long) {
            lds.poison();
            cx<T>* sm = lds.data();
            for (int ph = 0; ph < 3; ++ph)
                emu_threads(M::NT, [&](int tid) { multi_task_phase<T, PRESET>(ph, tid, tasks[t], a, sm, bk); });
            if (!lds.guards_ok() && g_emu_bad_block < 0) g_emu_bad_block = t;
        });
    return hipSuccess;
}

template <typename T, int PRESET>
hipError_t launch_multi_dev(const void* d, const void* bank, unsigned* work, int, hipStream_t)
{
    MultiDescriptor<T>* md = (MultiDescriptor<T>*)d;
    return launch_multi<T, PRESET>(&md->args, md->tasks, md->ntasks, bank, work, nullptr);
}

template <typename T, int PRESET>
int multi_dev_grid() { return 3; }

template <typename T, int PRESET>
MultiDesc multi_desc(const char* name, int order = 0)
{
    using M = MultiCfg<T, PRESET>;
    return MultiDesc{prec_of<T>(), PRESET, M::NT,
                     {M::vpw(0), M::vpw(1), M::vpw(2), M::vpw(3), M::vpw(4)},
                     sizeof(cx<T>) * (size_t)M::LDS, name, &launch_multi<T, PRESET>,
                     &launch_multi_dev<T, PRESET>, &multi_dev_grid<T, PRESET>, 0, nullptr, order};
}

}  // namespace fx
