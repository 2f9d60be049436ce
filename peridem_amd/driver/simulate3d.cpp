// simulate3d -- host driver of the MI355X explicit step.
//
// Same command line, config keys and output files as the reference driver
// (run/simulate3d.cpp:40-157 + run_timeloop, src/particle/timeloop.h:231-965):
//
//   simulate3d -c config/main.conf -i setup.h5 -o outdir
//
// It reads the setup file (P_%05u/{Pos, Vol, Connectivity, ...}, pairwise/*,
// wall/*; load_hdf5.h:41-181), applies the config overrides exactly as
// Timeloop/Contact/RectWall::apply_config do, hands the particles to the GPU
// library through the C-ABI of include/peridem_gpu.h, applies the reference's
// time switches between pd_step calls (timeloop.h:405-535) and writes
// plotinfo.h5, tc_%05u.h5, wall_%05u.h5 and run_time.h5 (timeloop.h:320-350,
// 799-964).  Files ending in .h5 are HDF5 (when built with HDF5); any other
// name is the plain "pdbin" container holding the same dataset paths.
//
// Extra options: --convert IN OUT copies a setup between containers;
// --print-config dumps the parsed config and exits.
#include <getopt.h>
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "peridem_gpu.h"
#include "store.h"

using std::string;
using std::vector;

// ---------------------------------------------------------------------------
// ConfigVal (src/read/read_config.h:14-461): keys, defaults and sentinels.
struct Config {
  unsigned timesteps = 8000, modulo = 100;
  double dt = 0.02 / 1e5;
  int is_parallel = 1, do_resume = 0, wall_resume = 0;
  unsigned resume_ind = 50;
  int save_file = 1, enable_fracture = 0, nl_bdry_only = 0;
  int allow_damping = 1, allow_friction = 1;
  double normal_stiffness = -1, damping_ratio = -1, friction_coefficient = -1, new_snot = -1;
  int self_contact = 1;
  double self_contact_rad = -1;
  int enable_cohesion = 0;
  double cohesion_scaling = 1;
  int enable_torque = 0, set_torque_timestep = -1, set_fracture_timestep = -1;
  int set_self_contact_timestep = -1, set_zero_wall_speed_timestep = -1, set_damping_timestep = -1;
  int set_movable_index = -1, set_movable_timestep = -1, set_stoppable_index = -1,
      set_stoppable_timestep = -1;
  int reset_partzero_y = 0;
  unsigned reset_partzero_y_timestep = 0;
  double wheel_rad = 0;
  int particle_w_fixed_velocity = -1, set_fix_velocity_timestep = -1, fix_velocity_angular = 0;
  double prescribed_angular_vel = -999;
  int fix_velocity_mean_x = 0;
  double prescribed_velocity_mean_x = -999;
  double wall_left = -999, wall_right = -999, wall_top = -999, wall_bottom = -999;
  double wall_x_min = -999, wall_y_min = -999, wall_z_min = -999, wall_x_max = -999,
         wall_y_max = -999, wall_z_max = -999;
  double speed_wall_left = 0, speed_wall_right = 0, speed_wall_top = 0, speed_wall_bottom = 0;
  double speed_wall_x_min = 0, speed_wall_y_min = 0, speed_wall_z_min = 0, speed_wall_x_max = 0,
         speed_wall_y_max = 0, speed_wall_z_max = 0;
  int brazil = 0, brazil_reset_wall_bottom_freq = -1, brazil_nonzero_wall_bottom_speed_freq = -1;
  double brazil_reset_wall_bottom_y = -999;
  int forcefield_type = 0;
  double forcefield_scaling = 1e9;
  int gradient_extforce = 0;
  unsigned extforce_maxstep = 0;
  int compute_wall_reaction = 0;
  string setup_filename = "data/hdf5/all.h5";
  string output_dir = "output/hdf5/";

  // read_file (read_config.h:170-435): `name = value`, whitespace removed,
  // '#' comments, unknown keys reported and ignored.
  void read_file(const string& filename) {
    std::ifstream f(filename);
    if (!f.is_open()) {
      std::cerr << "Couldn't open config file for reading.\n";
      return;
    }
    std::cout << "Reading config file: " << filename << std::endl;
    std::map<string, int*> I = {
        {"is_parallel", &is_parallel}, {"do_resume", &do_resume}, {"wall_resume", &wall_resume},
        {"save_file", &save_file}, {"enable_fracture", &enable_fracture},
        {"nl_bdry_only", &nl_bdry_only}, {"particle_w_fixed_velocity", &particle_w_fixed_velocity},
        {"set_fix_velocity_timestep", &set_fix_velocity_timestep},
        {"fix_velocity_angular", &fix_velocity_angular}, {"fix_velocity_mean_x", &fix_velocity_mean_x},
        {"allow_damping", &allow_damping}, {"allow_friction", &allow_friction},
        {"forcefield_type", &forcefield_type}, {"self_contact", &self_contact},
        {"enable_torque", &enable_torque}, {"set_torque_timestep", &set_torque_timestep},
        {"set_fracture_timestep", &set_fracture_timestep},
        {"set_self_contact_timestep", &set_self_contact_timestep},
        {"set_damping_timestep", &set_damping_timestep},
        {"set_zero_wall_speed_timestep", &set_zero_wall_speed_timestep},
        {"enable_cohesion", &enable_cohesion}, {"set_movable_index", &set_movable_index},
        {"set_movable_timestep", &set_movable_timestep}, {"set_stoppable_index", &set_stoppable_index},
        {"set_stoppable_timestep", &set_stoppable_timestep}, {"reset_partzero_y", &reset_partzero_y},
        {"brazil", &brazil}, {"brazil_reset_wall_bottom_freq", &brazil_reset_wall_bottom_freq},
        {"brazil_nonzero_wall_bottom_speed_freq", &brazil_nonzero_wall_bottom_speed_freq},
        {"gradient_extforce", &gradient_extforce}, {"compute_wall_reaction", &compute_wall_reaction}};
    std::map<string, unsigned*> Uu = {{"timesteps", &timesteps}, {"modulo", &modulo},
                                      {"resume_ind", &resume_ind},
                                      {"reset_partzero_y_timestep", &reset_partzero_y_timestep},
                                      {"extforce_maxstep", &extforce_maxstep}};
    std::map<string, double*> D = {
        {"dt", &dt}, {"prescribed_angular_vel", &prescribed_angular_vel},
        {"prescribed_velocity_mean_x", &prescribed_velocity_mean_x}, {"damping_ratio", &damping_ratio},
        {"friction_coefficient", &friction_coefficient}, {"normal_stiffness", &normal_stiffness},
        {"new_snot", &new_snot}, {"forcefield_scaling", &forcefield_scaling},
        {"self_contact_rad", &self_contact_rad}, {"cohesion_scaling", &cohesion_scaling},
        {"wheel_rad", &wheel_rad}, {"wall_left", &wall_left}, {"wall_right", &wall_right},
        {"wall_top", &wall_top}, {"wall_bottom", &wall_bottom}, {"speed_wall_left", &speed_wall_left},
        {"speed_wall_right", &speed_wall_right}, {"speed_wall_top", &speed_wall_top},
        {"speed_wall_bottom", &speed_wall_bottom},
        {"brazil_reset_wall_bottom_y", &brazil_reset_wall_bottom_y}, {"wall_x_min", &wall_x_min},
        {"wall_y_min", &wall_y_min}, {"wall_z_min", &wall_z_min}, {"wall_x_max", &wall_x_max},
        {"wall_y_max", &wall_y_max}, {"wall_z_max", &wall_z_max},
        {"speed_wall_x_min", &speed_wall_x_min}, {"speed_wall_y_min", &speed_wall_y_min},
        {"speed_wall_z_min", &speed_wall_z_min}, {"speed_wall_x_max", &speed_wall_x_max},
        {"speed_wall_y_max", &speed_wall_y_max}, {"speed_wall_z_max", &speed_wall_z_max}};
    string line;
    while (std::getline(f, line)) {
      string l;
      for (char ch : line)
        if (!std::isspace((unsigned char)ch)) l += ch;
      if (l.empty() || l[0] == '#') continue;
      size_t p = l.find('=');
      string name = l.substr(0, p), value = p == string::npos ? l : l.substr(p + 1);
      if (value.empty()) {
        std::cout << "[Error] Missing value for option: " << name << std::endl;
        throw std::runtime_error("missing value for option " + name);
      }
      if (I.count(name)) *I[name] = std::stoi(value);
      else if (Uu.count(name)) *Uu[name] = (unsigned)std::stoi(value);
      else if (D.count(name)) *D[name] = std::stod(value);
      else std::cerr << "[Error]: Wrong config name: " << name << " !!\n";
    }
    std::cout << "Reading config file: Done" << std::endl;
  }
};

// ---------------------------------------------------------------------------
// setup data (load_particles / load_contact / load_wall, load_hdf5.h:41-181)
struct Part {
  uint32_t n = 0;
  vector<double> pos, vol, disp, vel, acc, ext;
  vector<uint32_t> conn, bdry, clamped;
  double delta = 0, rho = 0, cnot = 0, snot = 0;
  int movable = 1, breakable = 1, stoppable = 1;
  uint32_t torque_axis = 0;
  double torque_val = 0;
};

static string pname(unsigned i) {
  char b[32];
  snprintf(b, sizeof b, "P_%05u", i);
  return b;
}

static vector<Part> load_particles(Store& s) {
  unsigned np = s.read_u32("total_particles_univ").at(0);
  std::cout << "Total particles: " << np << std::endl;
  vector<Part> P(np);
  for (unsigned i = 0; i < np; i++) {
    string g = pname(i) + "/";
    Part& p = P[i];
    p.pos = s.read_f64(g + "Pos");
    p.n = (uint32_t)(p.pos.size() / 3);
    p.vol = s.read_f64(g + "Vol");
    p.bdry = s.has(g + "bdry_nodes") ? s.read_u32(g + "bdry_nodes") : vector<uint32_t>();
    p.clamped = s.has(g + "clamped_nodes") ? s.read_u32(g + "clamped_nodes") : vector<uint32_t>();
    p.disp = s.read_f64(g + "disp");
    p.vel = s.read_f64(g + "vel");
    p.acc = s.read_f64(g + "acc");
    p.ext = s.read_f64(g + "extforce");
    p.conn = s.read_u32(g + "Connectivity");
    p.delta = s.read_f64(g + "delta").at(0);
    p.rho = s.read_f64(g + "rho").at(0);
    p.cnot = s.read_f64(g + "cnot").at(0);
    p.snot = s.read_f64(g + "snot").at(0);
    p.movable = s.read_i32(g + "movable").at(0);
    p.breakable = s.read_i32(g + "breakable").at(0);
    p.stoppable = s.read_i32(g + "stoppable").at(0);
    p.torque_axis = s.read_u32(g + "torque_axis").at(0);
    p.torque_val = s.read_f64(g + "torque_val").at(0);
    if (p.vol.size() != p.n || p.disp.size() != 3 * p.n || p.vel.size() != 3 * p.n ||
        p.acc.size() != 3 * p.n || p.ext.size() != 3 * p.n || p.conn.size() % 2)
      throw std::runtime_error("inconsistent dataset sizes in " + pname(i));
  }
  std::cout << "Done reading particles." << std::endl;
  return P;
}

struct Wall {
  bool allow = false;
  double lrtb[6] = {0, 0, 0, 0, 0, 0};   // x_min, y_min, z_min, x_max, y_max, z_max
  double speed[6] = {0, 0, 0, 0, 0, 0};
};

static void mkdirs(const string& d) {
  string cur;
  for (size_t i = 0; i <= d.size(); i++) {
    if (i == d.size() || d[i] == '/') {
      if (!cur.empty()) mkdir(cur.c_str(), 0755);
    }
    if (i < d.size()) cur += d[i];
  }
}

static string format_ms(long ms) {   // timeloop.h:46-58
  long s = ms / 1000, mn = s / 60, hr = mn / 60;
  std::stringstream ss;
  ss << hr << ":" << std::setfill('0') << std::setw(2) << mn % 60 << ":" << std::setfill('0')
     << std::setw(2) << s % 60;
  return ss.str();
}

static void check(pd_ctx* c, int rc, const char* what) {
  if (rc) throw std::runtime_error(string(what) + ": " + (c ? pd_last_error(c) : "error"));
}

static int run(Config& C) {
  auto in = open_store(C.setup_filename, false);
  vector<Part> P = load_particles(*in);
  // load_contact (load_hdf5.h:115-127)
  double Rc = in->read_f64("pairwise/contact_radius").at(0);
  double Kn = in->read_f64("pairwise/normal_stiffness").at(0);
  double mu = in->read_f64("pairwise/friction_coefficient").at(0);
  double zeta = in->read_f64("pairwise/damping_ratio").at(0);
  std::cout << "Done reading contact." << std::endl;
  // load_wall (load_hdf5.h:150-181)
  Wall W;
  W.allow = in->read_u32("wall/allow_wall").at(0) != 0;
  auto wi = in->read_f64("wall/geom_wall_info");
  if (wi.size() == 6)
    for (int k = 0; k < 6; k++) W.lrtb[k] = wi[k];
  std::cout << "Done reading wall size." << std::endl;
  in.reset();

  // Contact::apply_config (contact.h:156-182)
  if (C.damping_ratio != -1) zeta = C.damping_ratio;
  if (C.friction_coefficient != -1) mu = C.friction_coefficient;
  if (C.normal_stiffness != -1) Kn = C.normal_stiffness;
  // RectWall<3>::apply_config (contact.h:308-370): wall_top/right also set
  // z_max/y_max in 3-D; speeds always come from the config
  if (C.wall_top != -999) W.lrtb[5] = C.wall_top;
  if (C.wall_right != -999) W.lrtb[4] = C.wall_right;
  double* cw[6] = {&C.wall_x_min, &C.wall_y_min, &C.wall_z_min, &C.wall_x_max, &C.wall_y_max, &C.wall_z_max};
  for (int k = 0; k < 6; k++)
    if (*cw[k] != -999) W.lrtb[k] = *cw[k];
  double sp[6] = {C.speed_wall_x_min, C.speed_wall_y_min, C.speed_wall_z_min,
                  C.speed_wall_x_max, C.speed_wall_y_max, C.speed_wall_z_max};
  for (int k = 0; k < 6; k++) W.speed[k] = sp[k];

  // the brazil-nut schedule (timeloop.h:466-479) writes Wall.bottom and
  // Wall.speed_bottom, which a 3-D RectWall never reads (contact.h:221-238):
  // in 3-D it changes nothing
  if (C.brazil) std::cout << "brazil: acts on the 2-D wall members only, no effect in 3-D" << std::endl;
  // resume (timeloop.h:262-312)
  unsigned counter = 1, first_counter = 1, last_counter = C.timesteps / C.modulo;
  if (C.do_resume) {
    counter = C.resume_ind + 1;
    first_counter = counter;
    last_counter = C.resume_ind + C.timesteps / C.modulo;
    char b[32];
    snprintf(b, sizeof b, "/tc_%05u.h5", C.resume_ind);
    string fn = C.output_dir + b;
    std::cout << "Loading from file: " << fn << std::endl;
    auto r = open_store(fn, false);
    for (unsigned i = 0; i < P.size(); i++) {
      string g = pname(i) + "/";
      auto cp = r->read_f64(g + "CurrPos");
      for (size_t k = 0; k < cp.size(); k++) P[i].disp[k] = cp[k] - P[i].pos[k];
      P[i].vel = r->read_f64(g + "vel");
      P[i].acc = r->read_f64(g + "acc");
      if (C.enable_fracture) P[i].conn = r->read_u32(g + "Connectivity");
    }
    if (C.wall_resume) {
      snprintf(b, sizeof b, "/wall_%05u.h5", C.resume_ind);
      auto wr = open_store(C.output_dir + b, false);
      auto v = wr->read_f64("wall_info");
      for (int k = 0; k < 6 && k < (int)v.size(); k++) W.lrtb[k] = v[k];
    }
    std::cout << "Resuming from counter " << counter << std::endl;
  }
  if (C.new_snot != -1)
    for (auto& p : P) p.snot = C.new_snot;   // timeloop.h:314-318

  mkdirs(C.output_dir);
  {  // plotinfo.h5 (timeloop.h:320-350)
    auto pl = open_store(C.output_dir + "/plotinfo.h5", true);
    pl->write_u32("f_l_counter", {first_counter, last_counter}, {2, 1});
    pl->write_u32("dimension", {3}, {1, 1});
    pl->write_f64("dt", {C.dt}, {1, 1});
    pl->write_u32("modulo", {C.modulo}, {1, 1});
    pl->write_i32("wall/allow_wall", {W.allow ? 1 : 0}, {1, 1});
    pl->write_f64("wall/geom_wall_info", vector<double>(W.lrtb, W.lrtb + 6), {6, 1});
  }

  pd_run_desc r{};
  r.dt = C.dt;
  r.enable_fracture = C.enable_fracture;
  r.self_contact = C.self_contact;
  r.allow_damping = C.allow_damping;
  r.allow_friction = C.allow_friction;
  r.nl_bdry_only = C.nl_bdry_only;
  r.compute_wall_reaction = C.compute_wall_reaction;
  r.gradient_extforce = C.gradient_extforce;
  r.extforce_maxstep = C.extforce_maxstep;
  r.enable_torque = C.enable_torque;
  r.allow_wall = W.allow;
  r.contact_rad = Rc;
  r.normal_stiffness = Kn;
  r.friction_coefficient = mu;
  r.damping_ratio = zeta;
  for (int k = 0; k < 6; k++) {
    r.wall_lrtb[k] = W.lrtb[k];
    r.wall_speed[k] = W.speed[k];
  }
  r.first_step = 1;
  pd_ctx* ctx = nullptr;
  check(nullptr, pd_create(&ctx, &r), "pd_create");
  for (auto& p : P) {
    pd_particle_desc d{};
    d.nnodes = p.n;
    d.pos = p.pos.data();
    d.vol = p.vol.data();
    d.disp = p.disp.data();
    d.vel = p.vel.data();
    d.acc = p.acc.data();
    d.extforce = p.ext.data();
    d.nconn = p.conn.size() / 2;
    d.conn = p.conn.empty() ? nullptr : p.conn.data();
    d.nclamped = (uint32_t)p.clamped.size();
    d.clamped = p.clamped.empty() ? nullptr : p.clamped.data();
    d.nbdry = (uint32_t)p.bdry.size();
    d.bdry = p.bdry.empty() ? nullptr : p.bdry.data();
    d.delta = p.delta;
    d.rho = p.rho;
    d.cnot = p.cnot;
    d.snot = p.snot;
    d.movable = p.movable;
    d.breakable = p.breakable;
    d.stoppable = p.stoppable;
    d.torque_axis = p.torque_axis;
    d.torque_val = p.torque_val;
    check(ctx, pd_add_particle(ctx, &d, nullptr), "pd_add_particle");
  }
  check(ctx, pd_commit(ctx), "pd_commit");
  // prescribed velocity of one particle (timeloop.h:685-764), switched on at
  // set_fix_velocity_timestep
  const bool vc_particle = C.particle_w_fixed_velocity >= 0 && (unsigned)C.particle_w_fixed_velocity < P.size();
  if (vc_particle)
    check(ctx,
          pd_set_velocity_constraint(ctx, C.particle_w_fixed_velocity, C.fix_velocity_mean_x,
                                     C.prescribed_velocity_mean_x, C.fix_velocity_angular,
                                     C.prescribed_angular_vel),
          "pd_set_velocity_constraint");

  const unsigned cw_ = 12;
  time_t now = time(NULL);
  char sc[64];
  std::cout << "----------------------------------------------------------------------" << std::endl;
  std::cout << "Starting time loop (GPU): " << ctime(&now);
  snprintf(sc, sizeof sc, "ct[%u:%u]", first_counter, last_counter);
  std::cout << std::left << std::setw(cw_) << "t " << std::setw(cw_) << sc << std::setw(cw_ + 5)
            << "duration(s)" << std::setw(cw_) << "rem(h:m:s)" << std::setw(cw_) << "#contacts"
            << std::setw(cw_) << "time" << std::endl;

  // the steps at which the host acts before the step (timeloop.h:405-535)
  auto is_switch = [&](unsigned t) {
    for (int s : {C.set_torque_timestep, C.set_fix_velocity_timestep, C.set_fracture_timestep,
                  C.set_self_contact_timestep, C.set_damping_timestep, C.set_zero_wall_speed_timestep,
                  C.set_movable_timestep, C.set_stoppable_timestep})
      if (s != -1 && (unsigned)s == t) return true;
    return C.reset_partzero_y && C.reset_partzero_y_timestep == t;
  };
  auto apply_switches = [&](unsigned t) {
    if (C.set_torque_timestep != -1 && t == (unsigned)C.set_torque_timestep) {
      check(ctx, pd_set_flag(ctx, PD_FLAG_ENABLE_TORQUE, 1), "pd_set_flag");
      std::cout << "Setting torque to 1 on timestep " << t << std::endl;
    }
    if (C.set_fix_velocity_timestep != -1 && t == (unsigned)C.set_fix_velocity_timestep) {
      check(ctx, pd_set_flag(ctx, PD_FLAG_VELOCITY_CONSTRAINT, 1), "pd_set_flag");
      std::cout << "Setting velocity constraint to 1 on timestep " << t << std::endl;
    }
    if (C.set_fracture_timestep != -1 && t == (unsigned)C.set_fracture_timestep) {
      check(ctx, pd_set_flag(ctx, PD_FLAG_ENABLE_FRACTURE, 1), "pd_set_flag");
      std::cout << "Setting fracture to 1 on timestep " << t << std::endl;
    }
    if (C.set_self_contact_timestep != -1 && t == (unsigned)C.set_self_contact_timestep) {
      check(ctx, pd_set_flag(ctx, PD_FLAG_SELF_CONTACT, 1), "pd_set_flag");
      std::cout << "Setting self_contact to 1 on timestep " << t << std::endl;
    }
    if (C.set_damping_timestep != -1 && t == (unsigned)C.set_damping_timestep) {
      check(ctx, pd_set_flag(ctx, PD_FLAG_ALLOW_DAMPING, 1), "pd_set_flag");
      std::cout << "Setting damping to 1 on timestep " << t << std::endl;
    }
    if (C.set_zero_wall_speed_timestep != -1 && t == (unsigned)C.set_zero_wall_speed_timestep) {
      double z[6] = {0, 0, 0, 0, 0, 0};
      check(ctx, pd_set_wall_speed(ctx, z), "pd_set_wall_speed");
      std::cout << "Setting all wall speed to zero on timestep " << t << std::endl;
    }
    if (C.set_movable_index != -1 && (unsigned)C.set_movable_timestep == t) {
      check(ctx, pd_set_particle_flag(ctx, C.set_movable_index, PD_PFLAG_MOVABLE, 1), "movable");
      std::cout << "Setting particle " << C.set_movable_index << " to movable on timestep " << t << std::endl;
    }
    if (C.set_stoppable_index != -1 && (unsigned)C.set_stoppable_timestep == t) {
      check(ctx, pd_set_particle_flag(ctx, C.set_stoppable_index, PD_PFLAG_STOPPABLE, 1), "stoppable");
      std::cout << "Setting particle " << C.set_stoppable_index << " to stoppable on timestep " << t << std::endl;
    }
    // the wheel reset (timeloop.h:496-535): particle 0's reference positions
    // move in y; its bond vectors keep the setup positions
    if (C.reset_partzero_y && C.reset_partzero_y_timestep == t) {
      check(ctx, pd_reset_partzero_y(ctx, C.wheel_rad), "pd_reset_partzero_y");
      std::cout << "Resetting particle zero's mean y on timestep " << t << std::endl;
    }
  };

  vector<double> run_time, t_ind;
  vector<uint32_t> pw_vec;
  uint64_t pw_last = 0;
  auto block_start = std::chrono::system_clock::now();
  unsigned t = 1;
  while (t <= C.timesteps) {
    apply_switches(t);
    // advance through the next output step or up to the step before a switch
    unsigned end = t;
    while (end < C.timesteps && end % C.modulo != 0 && !is_switch(end + 1)) end++;
    check(ctx, pd_step(ctx, end - t + 1), "pd_step");
    t = end + 1;
    if (C.save_file && end % C.modulo == 0) {
      char b[32];
      snprintf(b, sizeof b, "%05u", counter);
      auto out = open_store(C.output_dir + "/tc_" + b + ".h5", true);
      for (unsigned i = 0; i < P.size(); i++) {
        string g = pname(i) + "/";
        size_t n3 = 3ull * P[i].n;
        vector<double> cp(n3), v(n3), a(n3), f(n3);
        check(ctx, pd_get_state(ctx, i, nullptr, cp.data(), v.data(), a.data(), f.data()), "pd_get_state");
        out->write_f64(g + "CurrPos", cp, {P[i].n, 3});
        out->write_f64(g + "vel", v, {P[i].n, 3});
        out->write_f64(g + "acc", a, {P[i].n, 3});
        out->write_f64(g + "force", f, {P[i].n, 3});   // force density, as the reference stores
        if (vc_particle && i == (unsigned)C.particle_w_fixed_velocity) {
          // zeros until the constraint first acts (the reference leaves the
          // array uninitialised until then)
          vector<double> af(n3);
          check(ctx, pd_get_applied_force(ctx, i, af.data()), "pd_get_applied_force");
          out->write_f64(g + "applied_force_density", af, {P[i].n, 3});
        }
        uint64_t m = 0;
        check(ctx, pd_get_connectivity(ctx, i, nullptr, 0, &m), "pd_get_connectivity");
        vector<uint32_t> conn(2 * m);
        check(ctx, pd_get_connectivity(ctx, i, conn.data(), m, &m), "pd_get_connectivity");
        out->write_u32(g + "Connectivity", conn, {m, 2});
      }
      out.reset();
      double lrtb[6], reac[18];
      check(ctx, pd_get_wall(ctx, 0, lrtb, reac), "pd_get_wall");
      auto wo = open_store(C.output_dir + "/wall_" + b + ".h5", true);
      wo->write_f64("wall_info", vector<double>(lrtb, lrtb + 6), {6, 1});
      if (C.compute_wall_reaction) wo->write_f64("reaction", vector<double>(reac, reac + 18), {6, 3});
      wo.reset();
      pd_counters cn;
      check(ctx, pd_get_counters(ctx, &cn), "pd_get_counters");
      uint64_t pw = cn.pair_computations - pw_last;
      pw_last = cn.pair_computations;
      auto stop = std::chrono::system_clock::now();
      long ms = (long)std::chrono::duration_cast<std::chrono::milliseconds>(stop - block_start).count();
      block_start = stop;
      // ETA from the mean of the last 10 blocks (timeloop.h:916-928)
      double avg = 0;
      int s0 = (int)run_time.size() - 10;
      if (s0 > 0)
        for (size_t k = s0; k < run_time.size(); k++) avg += run_time[k];
      avg /= 10;
      long rem = (long)((double)(last_counter - counter) * avg);
      time_t tn = time(NULL);
      char ts[80];
      strftime(ts, 80, "%F-%T", localtime(&tn));
      std::cout << std::left << std::setw(cw_) << end << std::setw(cw_) << counter << std::setw(cw_ + 5)
                << (double)ms / 1000 << std::setw(cw_) << format_ms(rem) << std::setw(cw_) << pw
                << std::setw(cw_) << ts << std::endl;
      run_time.push_back((double)ms);
      t_ind.push_back(end);
      pw_vec.push_back((uint32_t)pw);
      counter++;
    }
  }
  {  // run_time.h5 (timeloop.h:956-964; written to the output directory)
    auto rt = open_store(C.output_dir + "/run_time.h5", true);
    rt->write_f64("run_time", run_time, {run_time.size(), 1});
    rt->write_f64("t_ind", t_ind, {t_ind.size(), 1});
    rt->write_u32("pairwise_computations", pw_vec, {pw_vec.size(), 1});
  }
  pd_destroy(ctx);
  return 0;
}

int main(int argc, char** argv) {
  Config C;
  string config_file = "config/main.conf";
  bool print_config = false;
  if (argc >= 4 && string(argv[1]) == "--convert") {
    try {
      auto a = open_store(argv[2], false);
      auto b = open_store(argv[3], true);
      copy_store(*a, *b);
    } catch (std::exception& e) {
      std::cerr << "[Error] " << e.what() << std::endl;
      return 1;
    }
    return 0;
  }
  for (int i = 1; i < argc; i++) {
    if (string(argv[i]) == "--print-config") print_config = true;
    if (string(argv[i]) == "--features") {
#ifdef PD_WITH_HDF5
      std::cout << "hdf5 yes" << std::endl;
#else
      std::cout << "hdf5 no" << std::endl;
#endif
      return 0;
    }
  }

  int opt;
  static struct option lo[] = {{"print-config", no_argument, nullptr, 'p'},
                               {"features", no_argument, nullptr, 'f'},
                               {nullptr, 0, nullptr, 0}};
  while ((opt = getopt_long(argc, argv, "c:i:o:", lo, nullptr)) != -1) {
    switch (opt) {
      case 'c': config_file = optarg; break;
      case 'i': C.setup_filename = optarg; break;
      case 'o': C.output_dir = optarg; break;
      case 'p': break;
      default: break;
    }
  }
  try {
    C.read_file(config_file);
    if (C.modulo == 0) throw std::runtime_error("modulo must be positive");
    if (print_config) {
      std::cout << "timesteps " << C.timesteps << "\nmodulo " << C.modulo << "\ndt " << std::setprecision(17)
                << C.dt << "\nenable_fracture " << C.enable_fracture << "\nself_contact " << C.self_contact
                << "\nallow_damping " << C.allow_damping << "\nallow_friction " << C.allow_friction
                << "\ndamping_ratio " << C.damping_ratio << "\nfriction_coefficient "
                << C.friction_coefficient << "\nnormal_stiffness " << C.normal_stiffness
                << "\nspeed_wall_z_max " << C.speed_wall_z_max << "\nwall_z_max " << C.wall_z_max
                << "\ncompute_wall_reaction " << C.compute_wall_reaction << std::endl;
      return 0;
    }
    auto start = std::chrono::system_clock::now();
    int rc = run(C);
    auto stop = std::chrono::system_clock::now();
    std::cout << "Runtime: " << std::chrono::duration_cast<std::chrono::seconds>(stop - start).count()
              << "s" << std::endl;
    return rc;
  } catch (std::exception& e) {
    std::cerr << "[Error] " << e.what() << std::endl;
    return 1;
  }
}
