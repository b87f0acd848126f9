// ecsg_run -- drop-in for the reference's main.exe: the same command line,
// the same RSF inputs and outputs, the time loop run on the MI355X through
// libecsg_hip.so.
//
// Command line (main.cpp:31-145): key=value pairs, booleans y/n.
//   num_threads (accepted, unused), check_energy, radius (1..6; above 1 on a
//   single grid), dt [ms],
//   imode (1 ECI), verb, top_bc, num_grids, movie_dt, movie_nt,
//   movie_v{x,y,z}_flag; files buoy_<i>, s11_<i>, s13_<i> (i = 1..num_grids,
//   init_del_par_field.cpp:21-123), src (init_src.cpp:16-45), rec_in
//   (init_del_rec_trace.cpp:16-45); outputs rec_out (print.cpp:113-135),
//   energy_trace (print.cpp:31-51, with check_energy), res_trace (main.cpp:
//   284-302, with verb and num_grids > 1), movie_v?_<i> (init_del_movie.cpp).
//   A file key given no value names the file after the key, as Madagascar
//   does.  par=<file> reads further key=value pairs from a file.
// Additions of this driver:
//   precision=f64|f32   wavefield storage (default f64, the reference's build)
//   aniso=y             read the nine stiffnesses c11_<i> c22_<i> c33_<i>
//                       c23_<i> c13_<i> c12_<i> c44_<i> c55_<i> c66_<i>
//                       (code convention c44 = 2*C44) instead of s11/s13
//   vy_mirror_all=y     mirror vy across the y walls on every plane
//   sxz_solver=0|1|2|3  vx-sxz interface system: Jacobi, Chebyshev-accelerated,
//                       Jacobi by one workgroup per interface (up to 9728
//                       entries), or the direct block elimination
//   jacobi_tol=<cap>,<fac>  exit tolerance min(cap, fac * AvgAbs(B)) of the
//                       Jacobi sweeps (default: update_ghost_sxz.cpp:229-236's
//                       constants of the storage precision)
//   ii_vz_xend=y        imode=0: the working copy's form of the x-end vz ghost
//                       edge (update_ghost_vz.cpp:98-107) instead of the
//                       committed one the golden II traces were made with
//   graph=y|n           replay the time step from a HIP graph (default y)
//   cfl=<frac>          without dt: dt = frac x the energy-method CFL bound
//                       h_min / (vp sqrt 3) of the inputs (eci/SConstruct:
//                       107-114), reduced on the device (ecsg_cfl_dt)
//   sbp=y               print the summation-by-parts residuals of grid 1 after
//                       each stress update (TestSBP, test_sbp.cpp, as the
//                       committed main.cpp calls it) and write them to
//                       sbp_trace (n1 = steps, n2 = 7)
//   dry_run=y           read and check every input, print the mapped source and
//                       receiver positions, and stop before any device work
// Errors are reported on stderr with exit status 1 where the reference asserts.
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "ecsg_hip.h"
#include "rsf_io.hpp"

namespace {

struct Fail : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// ---------------------------------------------------------------------------
// command-line parameters (sf_init / iRSF par(0), main.cpp:31-33)
struct Params {
    std::map<std::string, std::string> kv;

    void parse(int argc, char **argv)
    {
        std::string text;
        for (int i = 1; i < argc; i++) {
            const std::string a = argv[i];
            const size_t eq = a.find('=');
            if (eq == std::string::npos || eq == 0) continue;
            const std::string k = a.substr(0, eq);
            if (k == "par") {
                std::ifstream f(rsf::unquote(a.substr(eq + 1)));
                if (!f) throw Fail("cannot open parameter file " + a.substr(eq + 1));
                std::string t((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
                rsf::parse_pairs(t, kv);
            } else {
                kv[k] = rsf::unquote(a.substr(eq + 1));
            }
        }
    }
    bool has(const std::string &k) const { return kv.count(k) != 0; }
    std::string str(const std::string &k, const std::string &def) const { return has(k) ? kv.at(k) : def; }
    long geti(const std::string &k, long def) const
    {
        if (!has(k)) return def;
        try {
            return std::stol(kv.at(k));
        } catch (...) {
            throw Fail("parameter " + k + "=" + kv.at(k) + " is not an integer");
        }
    }
    double getd(const std::string &k, double def) const
    {
        if (!has(k)) return def;
        try {
            return std::stod(kv.at(k));
        } catch (...) {
            throw Fail("parameter " + k + "=" + kv.at(k) + " is not a number");
        }
    }
    // Madagascar booleans: y/Y/1 true, n/N/0 false
    bool getb(const std::string &k, bool def) const
    {
        if (!has(k)) return def;
        const std::string &v = kv.at(k);
        if (v.empty()) throw Fail("parameter " + k + " is empty");
        const char c = v[0];
        if (c == 'y' || c == 'Y' || c == '1') return true;
        if (c == 'n' || c == 'N' || c == '0') return false;
        throw Fail("parameter " + k + "=" + v + " is not a boolean");
    }
    // file name of an RSF tag: the parameter's value, else the tag itself
    std::string file(const std::string &tag) const { return str(tag, tag); }
};

struct GridGeom {
    int n[3];
    double h[3];
    double origin[3];
};

// z-stacking search and rounding of init_src.cpp:33-45 and
// init_del_rec_trace.cpp:31-42 (strict >, C round)
void map_location(const std::vector<GridGeom> &G, double x, double y, double z, int out[4], const char *what)
{
    int g = 0;
    double tmp = z;
    while (tmp > G[g].h[2] * G[g].n[2]) {
        tmp -= G[g].h[2] * G[g].n[2];
        g++;
        if (g >= (int)G.size()) throw Fail(std::string(what) + " lies below the last grid");
    }
    out[0] = g;
    out[1] = (int)std::round((x - G[g].origin[0]) / G[g].h[0]);
    out[2] = (int)std::round((y - G[g].origin[1]) / G[g].h[1]);
    out[3] = (int)std::round(tmp / G[g].h[2]);
    for (int k = 0; k < 3; k++) {
        const int lim = G[g].n[k];
        if (out[1 + k] < 0 || out[1 + k] > lim) throw Fail(std::string(what) + " lies outside grid " + std::to_string(g + 1));
    }
}

// a primary-node cube of grid g, checked against the grid's node counts
std::vector<double> read_cube(const Params &P, const std::string &tag, const GridGeom &geom)
{
    rsf::Input in;
    in.open(P.file(tag));
    for (int d = 0; d < 3; d++)
        if (in.geti("n" + std::to_string(d + 1), -1) != geom.n[d] + 1)
            throw Fail(tag + ": n" + std::to_string(d + 1) + " differs from the buoyancy file's");
    return std::vector<double>(in.data.begin(), in.data.end());
}

class Ctx {
  public:
    ecsg_ctx *h = nullptr;
    ~Ctx()
    {
        if (h) ecsg_destroy(h);
    }
    void chk(int rc, const char *what) const
    {
        if (rc != ECSG_OK) throw Fail(std::string(what) + ": " + (h ? ecsg_last_error(h) : "no context"));
    }
};

// the field's own range (no halo) of wavefield sel: field.cpp:85-98
void interior(const std::vector<double> &buf, const int len[3], std::vector<float> &out)
{
    const int lx = len[0] - 2, ly = len[1] - 2, lz = len[2] - 2;
    out.resize((size_t)lx * ly * lz);
    size_t q = 0;
    for (int iz = 1; iz <= lz; iz++)
        for (int iy = 1; iy <= ly; iy++) {
            const double *row = buf.data() + ((size_t)iz * len[1] + iy) * len[0] + 1;
            for (int ix = 0; ix < lx; ix++) out[q++] = (float)row[ix];
        }
}

int run(int argc, char **argv)
{
    Params P;
    P.parse(argc, argv);

    const bool check_energy = P.getb("check_energy", false);
    const int radius = (int)P.geti("radius", 0);
    if (radius <= 0) throw Fail("radius must be given and positive");
    if (radius > 6) throw Fail("radius " + std::to_string(radius) + " is not supported (FD coefficients exist for 1..6)");
    double dt = P.getd("dt", -1.0);
    // without dt: a fraction of the energy-method CFL bound of the inputs
    // (eci/SConstruct:107-114), computed on the device once they are read
    const double cfl = P.getd("cfl", -1.0);
    if (!(dt > 0) && !(cfl > 0)) throw Fail("dt must be given and positive (or cfl=, a fraction of the CFL bound)");
    const int imode = (int)P.geti("imode", 1);
    const bool verb = P.getb("verb", false);
    const int top_bc = (int)P.geti("top_bc", 0);
    const int num_grids = (int)P.geti("num_grids", 1);
    if (num_grids < 1 || num_grids > ECSG_MAX_GRIDS) throw Fail("num_grids must be in 1.." + std::to_string(ECSG_MAX_GRIDS));
    const std::string prec = P.str("precision", "f64");
    if (prec != "f64" && prec != "f32") throw Fail("precision must be f64 or f32");
    const bool aniso = P.getb("aniso", false);
    const bool use_graph = P.getb("graph", true);
    const bool sbp = P.getb("sbp", false);

    const double movie_dt = P.getd("movie_dt", -1.0);
    if (!(movie_dt > 0)) throw Fail("movie_dt must be given and positive");
    const int movie_nt = (int)P.geti("movie_nt", -1);
    if (movie_nt <= 0) throw Fail("movie_nt must be given and positive");
    bool movie_flag[3] = {P.getb("movie_vx_flag", false), P.getb("movie_vy_flag", false), P.getb("movie_vz_flag", false)};

    // ----- grids and parameters (init_del_par_field.cpp:21-72) -----
    std::vector<GridGeom> G(num_grids);
    std::vector<rsf::Input> buoy_in(num_grids);
    for (int i = 0; i < num_grids; i++) {
        const std::string tag = "buoy_" + std::to_string(i + 1);
        buoy_in[i].open(P.file(tag));
        for (int d = 0; d < 3; d++) {
            const long nn = buoy_in[i].geti("n" + std::to_string(d + 1), -1) - 1;
            if (nn <= 0) throw Fail(tag + ": n" + std::to_string(d + 1) + " must be at least 2");
            G[i].n[d] = (int)nn;
            G[i].h[d] = buoy_in[i].getd("d" + std::to_string(d + 1), -1.0);
            if (!(G[i].h[d] > 0)) throw Fail(tag + ": d" + std::to_string(d + 1) + " must be positive");
            G[i].origin[d] = 0;
        }
        if (i >= 1) G[i].origin[2] = G[i - 1].origin[2] + G[i - 1].n[2] * G[i - 1].h[2];
    }

    // ----- parameters of every grid (init_del_par_field.cpp:74-150) -----
    struct GridParams {
        std::vector<double> buoy;
        std::vector<std::vector<double>> st;
    };
    std::vector<GridParams> gp(num_grids);
    for (int i = 0; i < num_grids; i++) {
        const std::string id = std::to_string(i + 1);
        gp[i].buoy.assign(buoy_in[i].data.begin(), buoy_in[i].data.end());
        buoy_in[i].data.clear();
        buoy_in[i].data.shrink_to_fit();
        gp[i].st.resize(9);
        if (aniso) {
            static const char *names[9] = {"c11_", "c22_", "c33_", "c23_", "c13_", "c12_", "c44_", "c55_", "c66_"};
            for (int k = 0; k < 9; k++) gp[i].st[k] = read_cube(P, names[k] + id, G[i]);
        } else {
            gp[i].st[0] = read_cube(P, "s11_" + id, G[i]);
            gp[i].st[4] = read_cube(P, "s13_" + id, G[i]);
        }
    }

    // ----- source (init_src.cpp:16-45) -----
    int src_loc[4];
    std::vector<float> src_trace;
    {
        rsf::Input src;
        src.open(P.file("src"));
        const long src_nt = src.geti("n1", 0);
        if (src_nt < 0 || src_nt > (long)src.data.size()) throw Fail("src: bad n1");
        map_location(G, src.getd("o2", 0.0), src.getd("o3", 0.0), src.getd("o4", 0.0), src_loc, "source");
        src_trace.assign(src.data.begin(), src.data.begin() + src_nt);
        std::cout << "== Source is placed at... " << std::endl;
        std::cout << "(grid#" << src_loc[0] << ", ix=" << src_loc[1] << ", iy=" << src_loc[2] << ", iz=" << src_loc[3] << ")" << std::endl;
    }

    // ----- receivers (init_del_rec_trace.cpp:16-55) -----
    int n_rec = 0;
    std::vector<int> rec_loc;
    {
        rsf::Input rec;
        rec.open(P.file("rec_in"));
        n_rec = (int)rec.geti("n2", 0);
        if (rec.geti("n1", 3) != 3) throw Fail("rec_in: n1 must be 3 (x y z per receiver)");
        if (n_rec < 0 || 3L * n_rec > (long)rec.data.size()) throw Fail("rec_in: bad n2");
        rec_loc.resize(4 * (size_t)n_rec);
        for (int r = 0; r < n_rec; r++)
            map_location(G, rec.data[3 * r], rec.data[3 * r + 1], rec.data[3 * r + 2], &rec_loc[4 * r], "receiver");
        std::cout << "== Record traces at... " << std::endl;
        for (int r = 0; r < n_rec; r++)
            std::cout << "(grid#" << rec_loc[4 * r] << ", ix=" << rec_loc[4 * r + 1] << ",  iy=" << rec_loc[4 * r + 2]
                      << ", iz=" << rec_loc[4 * r + 3] << ")" << std::endl;
    }
    const bool dry = P.getb("dry_run", false);
    ecsg_config cfg{};
    cfg.num_grids = num_grids;
    for (int i = 0; i < num_grids; i++)
        for (int d = 0; d < 3; d++) {
            cfg.n[i][d] = G[i].n[d];
            cfg.h[i][d] = G[i].h[d];
        }
    if (!(dt > 0)) {
        if (dry) throw Fail("dry_run needs dt (the CFL bound is computed on the device)");
        std::vector<const double *> bp(num_grids);
        std::vector<std::array<const double *, 9>> sp(num_grids);
        for (int i = 0; i < num_grids; i++) {
            bp[i] = gp[i].buoy.data();
            for (int k = 0; k < 9; k++) sp[i][k] = gp[i].st[k].empty() ? nullptr : gp[i].st[k].data();
        }
        // std::array<T, 9> has the layout of T[9]
        if (ecsg_cfl_dt(&cfg, bp.data(), reinterpret_cast<const double *const (*)[9]>(sp.data()), aniso ? 0 : 1, cfl,
                        &dt) != ECSG_OK)
            throw Fail("ecsg_cfl_dt failed");
        std::cout << "== dt = " << cfl << " x the CFL bound = " << dt << " ms" << std::endl;
    }
    // iterations niter = 0 .. ceil(movie_dt*movie_nt/dt) (main.cpp:165)
    const double last = std::ceil(movie_dt * movie_nt / dt);
    if (last > 2e9) throw Fail("too many time steps");
    const int nsteps = (int)last + 1;
    std::cout << "== " << nsteps << " time steps of " << dt << " ms" << std::endl;
    if (dry) return 0;  // inputs read and checked; no device work

    cfg.dt = dt;
    cfg.radius = radius;
    cfg.top_bc = top_bc;
    cfg.imode = imode;
    cfg.check_energy = check_energy ? 1 : 0;
    cfg.precision = prec == "f32" ? ECSG_F32 : ECSG_F64;
    cfg.max_steps = nsteps;
    cfg.vy_mirror_all = P.getb("vy_mirror_all", false) ? 1 : 0;
    cfg.sxz_solver = (int)P.geti("sxz_solver", 0);
    cfg.ii_vz_xend = P.getb("ii_vz_xend", false) ? 1 : 0;

    Ctx C;
    if (ecsg_create(&cfg, &C.h) != ECSG_OK || !C.h) throw Fail(std::string("ecsg_create: ") + (C.h ? ecsg_last_error(C.h) : "failed"));
    for (int i = 0; i < num_grids; i++) {
        const double *sp[9];
        for (int k = 0; k < 9; k++) sp[k] = aniso ? gp[i].st[k].data() : gp[i].st[k == 4 ? 4 : 0].data();
        C.chk(ecsg_set_params(C.h, i, gp[i].buoy.data(), sp, aniso ? 0 : 1), ("grid " + std::to_string(i + 1) + " parameters").c_str());
        gp[i] = GridParams();
    }
    if (P.has("jacobi_tol")) {
        const std::string v = P.str("jacobi_tol", "");
        const size_t comma = v.find(',');
        if (comma == std::string::npos) throw Fail("jacobi_tol must be <cap>,<fac>");
        double cap = 0, fac = 0;
        try {
            cap = std::stod(v.substr(0, comma));
            fac = std::stod(v.substr(comma + 1));
        } catch (...) {
            throw Fail("jacobi_tol=" + v + ": not two numbers");
        }
        C.chk(ecsg_set_jacobi_tol(C.h, cap, fac), "jacobi_tol");
    }
    C.chk(ecsg_set_source(C.h, src_loc[0], src_loc[1], src_loc[2], src_loc[3], 0, src_trace.data(), (int)src_trace.size()), "source");
    C.chk(ecsg_set_receivers(C.h, n_rec, rec_loc.data()), "receivers");

    // ----- movies (init_del_movie.cpp:4-47) -----
    static const char *vname[3] = {"vx", "vy", "vz"};
    std::vector<std::unique_ptr<rsf::Output>> movie(3 * num_grids);
    bool any_movie = false;
    for (int j = 0; j < 3; j++) {
        if (!movie_flag[j]) continue;
        any_movie = true;
        for (int i = 0; i < num_grids; i++) {
            int len[3];
            C.chk(ecsg_field_dims(C.h, i, j, len), "field dims");
            auto o = std::make_unique<rsf::Output>();
            o->path = P.file(std::string("movie_") + vname[j] + "_" + std::to_string(i + 1));
            o->put("n1", len[0] - 2);
            o->put("n2", len[1] - 2);
            o->put("n3", len[2] - 2);
            o->put("n4", movie_nt);
            for (int d = 0; d < 3; d++) o->putf("d" + std::to_string(d + 1), G[i].h[d]);
            o->putf("d4", movie_dt);
            for (int d = 0; d < 3; d++) o->putf("o" + std::to_string(d + 1), G[i].origin[d]);
            o->putf("o4", 0.0);
            o->putq("label1", "X"); o->putq("unit1", "m");
            o->putq("label2", "Y"); o->putq("unit2", "m");
            o->putq("label3", "Z"); o->putq("unit3", "m");
            o->putq("label4", "Time"); o->putq("unit4", "ms");
            o->open();
            movie[j * num_grids + i] = std::move(o);
        }
    }
    // frame steps: after iteration niter when niter*dt <= iframe*movie_dt < niter*dt + dt (main.cpp:253-259)
    std::vector<int> frame_step;
    for (int niter = 0, iframe = 0; niter < nsteps; niter++)
        if (iframe < movie_nt && (niter * dt <= iframe * movie_dt) && (iframe * movie_dt < niter * dt + dt)) {
            frame_step.push_back(niter);
            iframe++;
        }
    std::vector<double> fbuf;
    std::vector<float> fout;
    auto print_movie = [&]() {
        for (int j = 0; j < 3; j++) {
            if (!movie_flag[j]) continue;
            for (int i = 0; i < num_grids; i++) {
                int len[3];
                C.chk(ecsg_field_dims(C.h, i, j, len), "field dims");
                fbuf.resize((size_t)len[0] * len[1] * len[2]);
                C.chk(ecsg_get_field(C.h, i, j, fbuf.data()), "movie frame");
                interior(fbuf, len, fout);
                movie[j * num_grids + i]->append(fout.data(), fout.size());
            }
        }
    };

    // ----- time loop (main.cpp:165-261) -----
    const int ni = num_grids - 1;
    const bool residuals = verb && ni > 0;
    std::vector<std::vector<float>> res(9 * (size_t)std::max(ni, 0));  // res_vec[k][j] (init_del_res_trace.cpp)
    std::vector<float> sbp_tr;                                          // [7][steps]
    const auto t0 = std::chrono::steady_clock::now();
    if (!verb && !sbp) {
        // whole steps, graph-replayed, stopping at the movie frames
        int done = 0;
        size_t next = 0;
        while (done < nsteps) {
            int upto = nsteps;
            if (any_movie && next < frame_step.size()) upto = frame_step[next] + 1;
            if (upto > done) C.chk(ecsg_step_n(C.h, upto - done, use_graph ? 1 : 0), "time steps");
            done = upto;
            if (any_movie && next < frame_step.size() && frame_step[next] + 1 == done) {
                print_movie();
                next++;
            }
        }
    } else {
        // phase by phase, as main.cpp, with the residual diagnostics in between
        std::vector<double> rv(std::max(ni, 1));
        auto res_pair = [&](int which, int j0, int fv, int fs) {
            C.chk(ecsg_compute_res(C.h, which, rv.data()), "residual");
            for (int k = 0; k < ni; k++) {
                double a = 0, b = 0;
                C.chk(ecsg_field_avgabs(C.h, k + 1, fv, &a), "AvgAbs");
                C.chk(ecsg_field_avgabs(C.h, k + 1, fs, &b), "AvgAbs");
                res[9 * k + j0].push_back((float)rv[k]);
                res[9 * k + j0 + 1].push_back((float)a);
                res[9 * k + j0 + 2].push_back((float)b);
            }
        };
        size_t next = 0;
        for (int niter = 0; niter < nsteps; niter++) {
            if (verb) std::cerr << "niter = " << niter << std::endl;
            for (int i = 0; i < num_grids; i++) C.chk(ecsg_update_vel(C.h, i), "UpdateVel");
            C.chk(ecsg_record_receivers(C.h), "receivers");
            if (residuals) {
                res_pair(0, 0, ECSG_VX, ECSG_SXZ);  // compute_res.cpp:4-32
                res_pair(1, 3, ECSG_VY, ECSG_SYZ);  // compute_res.cpp:35-63
            }
            C.chk(ecsg_add_source(C.h), "source");
            for (int k = 0; k < ni; k++) C.chk(ecsg_update_ghost_vz(C.h, k), "UpdateGhostVZ");
            if (residuals) res_pair(2, 6, ECSG_VZ, ECSG_SZZ);  // compute_res.cpp:67-95
            if (check_energy) C.chk(ecsg_compute_energy(C.h, nullptr), "energy");
            for (int i = 0; i < num_grids; i++) C.chk(ecsg_update_stress(C.h, i), "UpdateStress");
            if (sbp) {
                double r7[7];
                C.chk(ecsg_sbp_residuals(C.h, 0, r7), "SBP residuals");
                std::cerr << "SBP_res1..7 =";
                for (int k = 0; k < 7; k++) {
                    std::cerr << " " << r7[k];
                    sbp_tr.push_back((float)r7[k]);
                }
                std::cerr << std::endl;
            }
            for (int k = 0; k < ni; k++) {
                C.chk(ecsg_update_ghost_sxz(C.h, k), "UpdateGhostSXZ");
                C.chk(ecsg_update_ghost_syz(C.h, k), "UpdateGhostSYZ");
            }
            C.chk(ecsg_end_step(C.h), "end of step");
            if (any_movie && next < frame_step.size() && frame_step[next] == niter) {
                print_movie();
                next++;
            }
        }
    }
    C.chk(ecsg_synchronize(C.h), "synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::cout << "Total elapsed time " << secs << " s." << std::endl;
    for (auto &m : movie)
        if (m) m->close();

    // ----- outputs (main.cpp:272-302) -----
    if (check_energy) {
        std::vector<double> e(nsteps);
        C.chk(ecsg_get_energy(C.h, e.data()), "energy trace");
        rsf::Output o;
        o.path = P.file("energy_trace");
        o.put("n1", nsteps);
        o.putf("o1", 0.0);
        o.putf("d1", dt);
        o.putq("label1", "Time"); o.putq("unit1", "ms");
        o.put("n2", 1); o.put("n3", 1); o.put("n4", 1);
        o.write(std::vector<float>(e.begin(), e.end()));
    }
    {
        std::vector<double> tr((size_t)std::max(n_rec, 1) * nsteps);
        C.chk(ecsg_get_traces(C.h, tr.data()), "receiver traces");
        rsf::Output o;
        o.path = P.file("rec_out");
        o.put("n1", nsteps);
        o.putf("o1", 0.0);
        o.putf("d1", dt);
        o.putq("label1", "Time"); o.putq("unit1", "ms");
        o.put("n2", n_rec);
        o.write(std::vector<float>(tr.begin(), tr.begin() + (size_t)n_rec * nsteps));
    }
    if (sbp) {
        rsf::Output o;
        o.path = P.file("sbp_trace");
        o.put("n1", nsteps);
        o.put("n2", 7);
        std::vector<float> t((size_t)7 * nsteps);
        for (int n = 0; n < nsteps; n++)
            for (int k = 0; k < 7; k++) t[(size_t)k * nsteps + n] = sbp_tr[(size_t)n * 7 + k];
        o.write(t);
    }
    if (residuals) {
        rsf::Output o;
        o.path = P.file("res_trace");
        o.put("n1", nsteps);
        o.put("n2", 9);
        o.put("n3", ni);
        std::vector<float> all;
        for (int k = 0; k < ni; k++)
            for (int j = 0; j < 9; j++) all.insert(all.end(), res[9 * k + j].begin(), res[9 * k + j].end());
        o.write(all);
    }
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::cerr << "ecsg_run: " << e.what() << std::endl;
        return 1;
    }
}
