// Run-time binding of the C ABI (include/lbdem.h). The host program loads one library that
// exports the ABI — the MI355X library (lbm-dem_amd/lib/liblbdem.so) or any other library with the
// same exports — with dlopen, so one executable can drive either implementation.
#pragma once

#include <string>

#include "../../include/lbdem.h"

namespace lbdem {

struct Api {
    int (*create)(const lbdem_params*, lbdem_ctx**) = nullptr;
    int (*destroy)(lbdem_ctx*) = nullptr;
    const char* (*last_error)(const lbdem_ctx*) = nullptr;
    int (*set_particles)(lbdem_ctx*, int, const int*, const double*, const double*) = nullptr;
    int (*set_objects)(lbdem_ctx*, int, const double*, const double*, const double*) = nullptr;
    int (*set_bodies)(lbdem_ctx*, int, const double*, const double*, const double*, const double*,
                      const double*) = nullptr;
    int (*init)(lbdem_ctx*) = nullptr;
    int (*dem_step)(lbdem_ctx*) = nullptr;
    int (*coupling_step)(lbdem_ctx*) = nullptr;
    int (*lb_step)(lbdem_ctx*) = nullptr;
    int (*fs_step)(lbdem_ctx*) = nullptr;
    int (*cycle)(lbdem_ctx*, int) = nullptr;
    int (*synchronize)(lbdem_ctx*) = nullptr;
    int (*get_info)(lbdem_ctx*, lbdem_info*) = nullptr;
    int (*get_types)(lbdem_ctx*, uint8_t*, uint8_t*, int32_t*) = nullptr;
    int (*get_fields)(lbdem_ctx*, double*, double*, double*) = nullptr;
    int (*get_particles)(lbdem_ctx*, double*) = nullptr;
    int (*get_objects)(lbdem_ctx*, double*) = nullptr;
    int (*get_walls)(lbdem_ctx*, double*) = nullptr;
    int (*get_bodies)(lbdem_ctx*, double*) = nullptr;
    int (*get_neighbors)(lbdem_ctx*, int, int*, int*) = nullptr;
    int (*total_mass)(lbdem_ctx*, double*) = nullptr;
    int (*set_restart_fluid)(lbdem_ctx*, const int*, const uint8_t*, int, const double*) = nullptr;
    int (*store_velocity)(lbdem_ctx*) = nullptr;
    int (*get_velocity)(lbdem_ctx*, double*, double*) = nullptr;
    int (*set_stress)(lbdem_ctx*, int) = nullptr;
    int (*get_stress)(lbdem_ctx*, double*, double*, double*) = nullptr;
    int (*get_particle_stress)(lbdem_ctx*, double*) = nullptr;

    void* handle = nullptr;
    std::string path;

    // Loads every entry point of the ABI from the shared library at `path`; throws
    // std::runtime_error naming the library and the missing symbol.
    static Api load(const std::string& path);
};

}  // namespace lbdem
