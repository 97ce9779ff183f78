/*
 * bdem_compat.cpp - the reference's Device_* API over the C ABI.
 *
 * Plain C++ (no HIP headers), compiled into libbdem.so.  Each function
 * forwards to its bdem_* counterpart on one process-wide context; on error it
 * prints "<function>: <message>" and exits, as the reference's error_check
 * path does (Device/DeviceInterface.cu:1891-1899).
 */
#include <cstdio>
#include <cstdlib>
#include "bdem_device_compat.h"

namespace {
bdem_ctx* g_ctx = nullptr;

void check(const char* where, int rc) {
  if (rc == BDEM_OK) return;
  std::fprintf(stderr, "%s: %s (status %d)\n", where, g_ctx ? bdem_last_error(g_ctx) : "no context", rc);
  std::exit(1);
}

bdem_ctx* ctx(const char* where) {
  if (!g_ctx) {
    std::fprintf(stderr, "%s: Device_Set_SimulationData has not been called\n", where);
    std::exit(1);
  }
  return g_ctx;
}
}  // namespace

void Device_Set_SimulationData(WOBJ* h_WorldOBJ, POBJ* h_ParticleOBJ, DOBJ* h_DynamicOBJ,
                               SimulationInfo* h_SimInfo, int h_Num_WorldOBJ, int h_Num_ParticleOBJ,
                               int h_Num_DynamicOBJ, InitConfig* h_PosConfig) {
  if (g_ctx) {
    bdem_destroy(g_ctx);
    g_ctx = nullptr;
  }
  const int dev = h_PosConfig ? h_PosConfig->use_device : 0;
  int rc = bdem_create(&g_ctx, dev);
  if (rc != BDEM_OK) {
    std::fprintf(stderr, "Device_Set_SimulationData: no usable HIP device %d (status %d)\n", dev, rc);
    std::exit(1);
  }
  check("Device_Set_SimulationData",
        bdem_set_simulation_data(g_ctx, h_WorldOBJ, h_ParticleOBJ, h_DynamicOBJ, h_SimInfo,
                                 h_Num_WorldOBJ, h_Num_ParticleOBJ, h_Num_DynamicOBJ, h_PosConfig));
}

void Device_Set_Pos(float3* h_pos, Quaterion* h_Ornt, uint* h_ptype, int* h_pid) {
  check("Device_Set_Pos", bdem_set_pos(ctx("Device_Set_Pos"), h_pos, h_Ornt, h_ptype, h_pid));
}

void Device_Set_Pos_Velocity(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             uint* h_ptype, int* h_pid) {
  check("Device_Set_Pos_Velocity", bdem_set_pos_velocity(ctx("Device_Set_Pos_Velocity"), h_pos,
                                                         h_Ornt, h_vel, h_Avel, h_ptype, h_pid));
}

void Device_Set_System_State(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             float3* h_acc, uint* h_ptype, int* h_pid) {
  check("Device_Set_System_State", bdem_set_system_state(ctx("Device_Set_System_State"), h_pos,
                                                         h_Ornt, h_vel, h_Avel, h_acc, h_ptype, h_pid));
}

void Device_Set_Tallys(float* h_Tallys) {
  check("Device_Set_Tallys", bdem_set_tallys(ctx("Device_Set_Tallys"), h_Tallys));
}

void Device_Get_System_State(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             float3* h_acc, uint* h_ptype, int* h_pid) {
  check("Device_Get_System_State", bdem_get_system_state(ctx("Device_Get_System_State"), h_pos,
                                                         h_Ornt, h_vel, h_Avel, h_acc, h_ptype, h_pid));
}

void Device_Get_P_Positions(float3* Pos, Quaterion* Quart_ornt, uint* h_ptype, int* h_pid) {
  check("Device_Get_P_Positions",
        bdem_get_positions(ctx("Device_Get_P_Positions"), Pos, Quart_ornt, h_ptype, h_pid));
}

void Device_Get_P_Velocity(float3* H_Vel, float3* H_Rvel, int Flag, uint* h_ptype, int* h_pid) {
  check("Device_Get_P_Velocity",
        bdem_get_velocity(ctx("Device_Get_P_Velocity"), H_Vel, H_Rvel, Flag, h_ptype, h_pid));
}

void Device_Get_W_Forces(float3* WallForces, float3* WallContact_Point) {
  check("Device_Get_W_Forces",
        bdem_get_wall_forces(ctx("Device_Get_W_Forces"), WallForces, WallContact_Point));
}

void Device_Get_Energy(float* Energy) {
  check("Device_Get_Energy", bdem_get_energy(ctx("Device_Get_Energy"), Energy));
}

void Device_Update_DObjectPostions(DOBJ* h_DynamicOBJ) {
  check("Device_Update_DObjectPostions",
        bdem_update_dobject_positions(ctx("Device_Update_DObjectPostions"), h_DynamicOBJ));
}

void Device_Update_WorldObjectPostions(WOBJ* h_WorldOBJ) {
  check("Device_Update_WorldObjectPostions",
        bdem_update_world_object_positions(ctx("Device_Update_WorldObjectPostions"), h_WorldOBJ));
}

void Device_OpenHatch(WOBJ* h_WorldOBJ, int h_Num_WorldOBJ) {
  check("Device_OpenHatch", bdem_open_hatch(ctx("Device_OpenHatch"), h_WorldOBJ, h_Num_WorldOBJ));
  std::printf("Hatch Opened\n");
}

void Device_Clean() {
  if (g_ctx) {
    bdem_destroy(g_ctx);
    g_ctx = nullptr;
  }
}

void Device_Reset_Energy(int num_values) {
  check("Device_Reset_Energy", bdem_reset_energy(ctx("Device_Reset_Energy"), num_values));
}

void drum_rotation(bool cylinder_rotation) {
  check("drum_rotation", bdem_drum_rotation(ctx("drum_rotation"), cylinder_rotation ? 1 : 0));
}

void Start_Bin() { check("Start_Bin", bdem_start_bin(ctx("Start_Bin"))); }
void Device_AddForce() { check("Device_AddForce", bdem_add_force(ctx("Device_AddForce"))); }
void Device_SubForce() { check("Device_SubForce", bdem_sub_force(ctx("Device_SubForce"))); }
void Device_ZeroForce() { check("Device_ZeroForce", bdem_zero_force(ctx("Device_ZeroForce"))); }

void Device_DEM_UpdateSim() {
  check("Device_DEM_UpdateSim", bdem_update_sim(ctx("Device_DEM_UpdateSim")));
}

bdem_ctx* Device_Compat_Context() { return g_ctx; }
