/*
 * compat_run.cpp - a caller written the way the reference's main.cpp drives
 * its device layer (main.cpp:942-946 setup, :242 UpdateSim loop,
 * Input/FileIO.h:615-645 state read-back), linked against libbdem.so
 * through the reference-named API of bdem_device_compat.h only.
 *
 *   compat_run <in.dump> <out.dump>
 *
 * in.dump: "BDEMDUMP", int32 {num_wobj, num_pobj, num_dobj, steps,
 * open_hatch_at}, then SimulationInfo, InitConfig, WOBJ[num_wobj],
 * POBJ[num_pobj], DOBJ[num_dobj] as raw structs, then the state of
 * Num_Particles particles as arrays: pos float3, orientation Quaterion,
 * velocity float3, angular velocity float3, type uint32, P_ID int32.
 * The program uploads the state (Device_Set_System_State), runs `steps`
 * Device_DEM_UpdateSim calls (Device_OpenHatch before step open_hatch_at
 * when it is >= 0), reads the state back (Device_Get_System_State) and
 * writes the same six arrays to out.dump, then calls Device_Clean.  Errors
 * exit(1), as the reference's error_check path does.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include "bdem_device_compat.h"

namespace {
template <class T>
void read_n(FILE* f, T* p, size_t n) {
  if (n && std::fread(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "compat_run: short input\n");
    std::exit(1);
  }
}
template <class T>
void write_n(FILE* f, const T* p, size_t n) {
  if (n && std::fwrite(p, sizeof(T), n, f) != n) {
    std::fprintf(stderr, "compat_run: write failed\n");
    std::exit(1);
  }
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: compat_run <in.dump> <out.dump>\n");
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) {
    std::fprintf(stderr, "compat_run: cannot open %s\n", argv[1]);
    return 1;
  }
  char magic[8];
  read_n(in, magic, 8);
  if (std::memcmp(magic, "BDEMDUMP", 8) != 0) {
    std::fprintf(stderr, "compat_run: not a dump file\n");
    return 1;
  }
  int hdr[5];
  read_n(in, hdr, 5);
  const int nw = hdr[0], np = hdr[1], nd = hdr[2], steps = hdr[3], hatch_at = hdr[4];
  if (nw < 0 || nw > 8 || np < 1 || np > 8 || nd < 0 || nd > 32 || steps < 0) {
    std::fprintf(stderr, "compat_run: bad header\n");
    return 1;
  }
  SimulationInfo sim;
  InitConfig init;
  read_n(in, &sim, 1);
  read_n(in, &init, 1);
  std::vector<WOBJ> wobj(nw > 0 ? nw : 1);
  std::vector<POBJ> pobj(np);
  std::vector<DOBJ> dobj(nd > 0 ? nd : 1);
  read_n(in, wobj.data(), nw);
  read_n(in, pobj.data(), np);
  read_n(in, dobj.data(), nd);
  const size_t n = sim.Num_Particles > 0 ? size_t(sim.Num_Particles) : 0;
  std::vector<float3> pos(n), vel(n), avel(n), acc(n);
  std::vector<Quaterion> q(n);
  std::vector<uint> ptype(n);
  std::vector<int> pid(n);
  read_n(in, pos.data(), n);
  read_n(in, q.data(), n);
  read_n(in, vel.data(), n);
  read_n(in, avel.data(), n);
  read_n(in, ptype.data(), n);
  read_n(in, pid.data(), n);
  std::fclose(in);

  Device_Set_SimulationData(wobj.data(), pobj.data(), nd > 0 ? dobj.data() : nullptr, &sim, nw, np, nd, &init);
  Device_Set_System_State(pos.data(), q.data(), vel.data(), avel.data(), acc.data(), ptype.data(), pid.data());
  for (int s = 0; s < steps; ++s) {
    if (s == hatch_at) Device_OpenHatch(wobj.data(), nw);
    Device_DEM_UpdateSim();
  }
  Device_Get_System_State(pos.data(), q.data(), vel.data(), avel.data(), acc.data(), ptype.data(), pid.data());
  Device_Clean();

  FILE* out = std::fopen(argv[2], "wb");
  if (!out) {
    std::fprintf(stderr, "compat_run: cannot write %s\n", argv[2]);
    return 1;
  }
  write_n(out, pos.data(), n);
  write_n(out, q.data(), n);
  write_n(out, vel.data(), n);
  write_n(out, avel.data(), n);
  write_n(out, ptype.data(), n);
  write_n(out, pid.data(), n);
  std::fclose(out);
  return 0;
}
