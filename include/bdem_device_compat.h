/*
 * bdem_device_compat.h - the reference's device-layer API, implemented by
 * libbdem.so on top of the C ABI (bdem_abi.h).
 *
 * Declares the C++ free functions of Device/DeviceInterface.h:16-95 with the
 * same names, parameter types and meaning, so a reference caller (KDevice,
 * Interface/KDevice.h:17-99, and main.cpp) links against libbdem.so instead
 * of DeviceInterface.o.  The structs carry the reference's type names
 * (float3, Quaterion, SimulationInfo, POBJ, WOBJ, DOBJ, InitConfig) so the
 * C++ symbol names match; each one is layout-identical to its bdem_*
 * counterpart.  A program that also includes HIP or CUDA vector headers
 * must not include this file in the same translation unit (both define
 * float3); include bdem_abi.h there instead.
 *
 * Error behaviour follows the reference: a failing call prints the message
 * and calls exit(1) (Device/DeviceInterface.cu:1891-1899).  The functions
 * work on one process-wide context created by Device_Set_SimulationData on
 * device InitConfig::use_device; Device_Compat_Context() exposes it for
 * callers that want the status-returning C ABI as well.
 */
#ifndef BDEM_DEVICE_COMPAT_H_
#define BDEM_DEVICE_COMPAT_H_

#ifndef __cplusplus
#error "bdem_device_compat.h is the C++ compatibility API; C callers use bdem_abi.h"
#endif

#include "bdem_abi.h"

typedef unsigned int uint;

struct float3 : bdem_float3 {};
struct int3 : bdem_int3 {};
struct Quaterion : bdem_Quaterion {};
struct SimulationInfo : bdem_SimulationInfo {};
struct InitConfig : bdem_InitConfig {};
struct POBJ : bdem_POBJ {};
struct WOBJ : bdem_WOBJ {};
struct DOBJ : bdem_DOBJ {};

static_assert(sizeof(float3) == sizeof(bdem_float3), "float3 layout");
static_assert(sizeof(Quaterion) == sizeof(bdem_Quaterion), "Quaterion layout");
static_assert(sizeof(SimulationInfo) == sizeof(bdem_SimulationInfo), "SimulationInfo layout");
static_assert(sizeof(InitConfig) == sizeof(bdem_InitConfig), "InitConfig layout");
static_assert(sizeof(POBJ) == sizeof(bdem_POBJ), "POBJ layout");
static_assert(sizeof(WOBJ) == sizeof(bdem_WOBJ), "WOBJ layout");
static_assert(sizeof(DOBJ) == sizeof(bdem_DOBJ), "DOBJ layout");

/* Set methods, Device/DeviceInterface.h:16-36 */
void Device_Set_SimulationData(WOBJ* h_WorldOBJ, POBJ* h_ParticleOBJ, DOBJ* h_DynamicOBJ,
                               SimulationInfo* h_SimInfo, int h_Num_WorldOBJ, int h_Num_ParticleOBJ,
                               int h_Num_DynamicOBJ, InitConfig* h_PosConfig);
void Device_Set_Pos(float3* h_pos, Quaterion* h_Ornt, uint* h_ptype, int* h_pid);
void Device_Set_Pos_Velocity(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             uint* h_ptype, int* h_pid);
void Device_Set_System_State(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             float3* h_acc, uint* h_ptype, int* h_pid);
void Device_Set_Tallys(float* h_Tallys);

/* Get methods, Device/DeviceInterface.h:42-63 */
void Device_Get_System_State(float3* h_pos, Quaterion* h_Ornt, float3* h_vel, float3* h_Avel,
                             float3* h_acc, uint* h_ptype, int* h_pid);
void Device_Get_P_Positions(float3* Pos, Quaterion* Quart_ornt, uint* h_ptype, int* h_pid);
void Device_Get_P_Velocity(float3* H_Vel, float3* H_Rvel, int Flag, uint* h_ptype, int* h_pid);
void Device_Get_W_Forces(float3* WallForces, float3* WallContact_Point);
void Device_Get_Energy(float* Energy);

/* Action methods, Device/DeviceInterface.h:72-85 */
void Device_Update_DObjectPostions(DOBJ* h_DynamicOBJ);
void Device_Update_WorldObjectPostions(WOBJ* h_WorldOBJ);
void Device_OpenHatch(WOBJ* h_WorldOBJ, int h_Num_WorldOBJ);
void Device_Clean();
void Device_Reset_Energy(int num_values);
void drum_rotation(bool cylinder_rotation);
void Start_Bin();
void Device_AddForce();
void Device_SubForce();
void Device_ZeroForce();

/* The step, Device/DeviceInterface.h:95 */
void Device_DEM_UpdateSim();

/* No reference counterpart: the context behind the functions above (NULL
 * before Device_Set_SimulationData). */
bdem_ctx* Device_Compat_Context();

#endif /* BDEM_DEVICE_COMPAT_H_ */
