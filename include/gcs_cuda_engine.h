// gcs_cuda_engine.h -- the reference's GPU engine interface, served by the
// MI355X library.
//
// Declares exactly the ten C++-linkage functions of the reference's
// cuda_engine.h (cuda_engine.h:5-32) over the reference's record types
// (types.h:5-18), so a driver written against the reference (its main.cpp
// Bake loop, main.cpp:119-297) links against libgcs_amd.so unchanged.  The
// definitions (gcs_cuda_engine.cpp) keep one process-wide engine, as the
// reference keeps its device state in file-scope globals
// (cuda_engine.cu:17-44).
#pragma once
#include <cstdint>

// types.h:5-12
struct Star {
    float x, y, z;
    float vx, vy, vz;
    float mass;
    uint8_t type;
    uint8_t r, g, b;
};

// types.h:14-18
struct PlaybackStar {
    float x, y, z;
    uint8_t r, g, b;
    uint8_t padding;
};

static_assert(sizeof(Star) == 32, "Star is 32 bytes");
static_assert(sizeof(PlaybackStar) == 16, "PlaybackStar is 16 bytes");

// cuda_engine.h:5 -- creates the engine for num_stars particles on device 0;
// false when there is no device or an allocation fails (cuda_engine.cu:219-285).
bool cuda_init(int num_stars);
// cuda_engine.h:6 -- frees the engine (cuda_engine.cu:303-340).
void cuda_cleanup();
// cuda_engine.h:9-12 -- page-locked host buffers (cuda_engine.cu:287-301).
void cuda_allocate_pinned_stars(Star** stars, int num_stars);
void cuda_free_pinned_stars(Star* stars);
void cuda_allocate_pinned_playback(PlaybackStar** buf, int size);
void cuda_free_pinned_playback(PlaybackStar* buf);
// cuda_engine.h:15 -- copies the initial Star records to the device
// (cuda_engine.cu:342-352).
void cuda_upload_initial_stars(const Star* host_stars, int num_stars);
// cuda_engine.h:18-19 -- OpenGL buffer registration.  The MI355X is headless:
// these record nothing, and a step asked for interop writes host_stars as the
// non-interop path does (cuda_engine.cu:354-363, :427-449).
void cuda_register_vbo(unsigned int vbo_id);
void cuda_unregister_vbo();
// cuda_engine.h:20 -- COM of the root of the last tree built
// (cuda_engine.cu:365-376); zeros before the first step.
void cuda_get_center_of_mass(float* cx, float* cy, float* cz);
// cuda_engine.h:23-32 -- one step with the reference's semantics
// (cuda_engine.cu:379-451): cube bounds [global_min, global_max], Morton sort,
// Karras tree, the reference's traversal rule (GCS_MAC_REFERENCE), kick-drift
// and speed colours; host_stars receives the state in device (sorted) order.
// Returns dt.
float cuda_physics_step(Star* host_stars, int num_stars, float G, float epsilon_sq, float dt,
                        float global_min, float global_max, bool use_interop);
