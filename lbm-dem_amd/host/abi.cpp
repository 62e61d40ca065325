#include "abi.hpp"

#include <dlfcn.h>

#include <stdexcept>

namespace lbdem {

template <typename F>
static void bind(void* h, const std::string& path, const char* name, F& fn) {
    void* s = dlsym(h, name);
    if (!s) throw std::runtime_error(path + ": missing entry point " + name);
    fn = reinterpret_cast<F>(s);
}

Api Api::load(const std::string& path) {
    Api a;
    a.handle = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!a.handle) throw std::runtime_error("cannot load " + path + ": " + dlerror());
    a.path = path;
    bind(a.handle, path, "lbdem_create", a.create);
    bind(a.handle, path, "lbdem_destroy", a.destroy);
    bind(a.handle, path, "lbdem_last_error", a.last_error);
    bind(a.handle, path, "lbdem_set_particles", a.set_particles);
    bind(a.handle, path, "lbdem_set_objects", a.set_objects);
    bind(a.handle, path, "lbdem_set_bodies", a.set_bodies);
    bind(a.handle, path, "lbdem_init", a.init);
    bind(a.handle, path, "lbdem_dem_step", a.dem_step);
    bind(a.handle, path, "lbdem_coupling_step", a.coupling_step);
    bind(a.handle, path, "lbdem_lb_step", a.lb_step);
    bind(a.handle, path, "lbdem_fs_step", a.fs_step);
    bind(a.handle, path, "lbdem_cycle", a.cycle);
    bind(a.handle, path, "lbdem_synchronize", a.synchronize);
    bind(a.handle, path, "lbdem_get_info", a.get_info);
    bind(a.handle, path, "lbdem_get_types", a.get_types);
    bind(a.handle, path, "lbdem_get_fields", a.get_fields);
    bind(a.handle, path, "lbdem_get_particles", a.get_particles);
    bind(a.handle, path, "lbdem_get_objects", a.get_objects);
    bind(a.handle, path, "lbdem_get_walls", a.get_walls);
    bind(a.handle, path, "lbdem_get_bodies", a.get_bodies);
    bind(a.handle, path, "lbdem_get_neighbors", a.get_neighbors);
    bind(a.handle, path, "lbdem_total_mass", a.total_mass);
    bind(a.handle, path, "lbdem_set_restart_fluid", a.set_restart_fluid);
    bind(a.handle, path, "lbdem_store_velocity", a.store_velocity);
    bind(a.handle, path, "lbdem_set_stress", a.set_stress);
    bind(a.handle, path, "lbdem_get_stress", a.get_stress);
    bind(a.handle, path, "lbdem_get_particle_stress", a.get_particle_stress);
    bind(a.handle, path, "lbdem_get_velocity", a.get_velocity);
    return a;
}

}  // namespace lbdem
