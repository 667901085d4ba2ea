// C++ host layer over the C ABI: setup-file front-end, FieldSolver plugins that keep their
// state on the GPU, and the FieldSimulation / Process driver.
#include "mpulse_host.h"
#include "mpx_hdf5.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <mutex>
#include <thread>

namespace mpulse_hip {

// ---------------------------------------------------------------------------------------
// Rebuildable: the token grammar of src/rebuild.cpp:15-116 and src/parameter.t:1-27
// ---------------------------------------------------------------------------------------

static void skip_comment_lines(std::istream &in, std::string &tok)
{
    while (tok == "//") {
        std::string rest;
        std::getline(in, rest);
        if (!(in >> tok)) {
            tok.clear();
            return;
        }
    }
}

std::string Rebuildable::Rebuild(std::istream &in)
{
    ParamMap pm;
    MakeParamMap(pm);
    std::string tok;
    if (!(in >> tok)) return "";
    while (tok != "}" && !tok.empty()) {
        skip_comment_lines(in, tok);
        if (tok.empty() || tok == "}") break;
        auto it = pm.find(tok);
        if (it == pm.end()) {
            std::string name = tok;
            if (!(in >> tok)) {
                tok.clear();
                break;
            }
            if (tok == "{") {
                warn("Unknown Task " + name + ". Skipping... ");
                int level = 0;
                do {
                    if (!(in >> tok)) {
                        tok.clear();
                        break;
                    }
                    if (tok == "{") level++;
                    if (tok == "}") level--;
                } while (tok != "}" || level >= 0);
                if (tok.empty()) {
                    warn("Unexpected end of file.");
                    break;
                }
                if (!(in >> tok)) tok.clear();
            } else {
                // the value is not consumed: it is read as the next key, as in the reference
                warn("Unknown Parameter " + name);
            }
            continue;
        }
        Param &p = it->second;
        if (p.kind == Param::INT) {
            in >> *p.ip;
            if (!(in >> tok)) tok.clear();
        } else if (p.kind == Param::DOUBLE) {
            in >> *p.dp;
            if (!(in >> tok)) tok.clear();
        } else if (p.kind == Param::STRING) {
            in >> *p.sp;
            if (!(in >> tok)) tok.clear();
        } else {
            Rebuildable *child = p.make();
            p.sink(p.owner, child);
            if (!(in >> tok)) {
                tok.clear();
                break;
            }
            if (tok == "{") tok = child->Rebuild(in);  // braces after an object key are optional
        }
    }
    if (tok == "}") {
        if (!(in >> tok)) tok.clear();
    }
    skip_comment_lines(in, tok);
    return tok;
}

template <class T>
static Rebuildable *make_new()
{
    return new T();
}

template <class Owner, class Base, class T>
static Param pobject(Owner *owner, void (*sink)(void *, Rebuildable *))
{
    Param q;
    q.kind = Param::OBJECT;
    q.make = &make_new<T>;
    q.sink = sink;
    q.owner = owner;
    return q;
}

// ---------------------------------------------------------------------------------------
// parameter maps (keys and defaults of the reference's MakeParamMap)
// ---------------------------------------------------------------------------------------

void Globals::MakeParamMap(ParamMap &pm)
{
    pm["grid-x"] = pint(&GridX, 128);
    pm["grid-y"] = pint(&GridY, 128);
    pm["grid-z"] = pint(&GridZ, 128);
    pm["grid-dx"] = pdouble(&GridDX, 0.1);
    pm["grid-dy"] = pdouble(&GridDY, 0.1);
    pm["grid-dz"] = pdouble(&GridDZ, 0.1);
    pm["dt"] = pdouble(&DT, 1e-2);
    pm["T-total"] = pint(&TotalTime, 100000);
}

void CPMLBorder::MakeParamMap(ParamMap &pm)
{
    pm["d"] = pint(&thickness, 8);
    pm["kappaMax"] = pdouble(&kappaMax, 15);
    pm["aMax"] = pdouble(&aMax, 0.25);
    pm["sigmaMax"] = pdouble(&sigmaMax, 3);
}

mpx_factory CPMLBorder::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, kind());
    f.d = thickness;
    f.kappa_max = kappaMax;
    f.a_max = aMax;
    f.sigma_max = sigmaMax;
    return f;
}

void IncidentSource::MakeParamMap(ParamMap &pm) { pm["d"] = pint(&distance, 15); }

void ShortPulseInject::MakeParamMap(ParamMap &pm)
{
    IncidentSource::MakeParamMap(pm);
    pm["length"] = pdouble(&length, 1.);
    pm["width"] = pdouble(&width, 1.);
    pm["om0"] = pdouble(&om0, 2 * 3.14159265358979323846);
    pm["TShift"] = pdouble(&TShift, 0);
    pm["ZShift"] = pdouble(&ZShift, 0);
    pm["Phase"] = pdouble(&Phase, 0);
    pm["amp"] = pdouble(&amp, 1);
    pm["eps"] = pdouble(&eps, 1);
}

mpx_factory ShortPulseInject::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_SHORT_PULSE);
    f.d = distance;
    f.length = length;
    f.width = width;
    f.om0 = om0;
    f.tshift = TShift;
    f.zshift = ZShift;
    f.phase = Phase;
    f.amp = amp;
    f.eps = eps;
    return f;
}

void GaussInject::MakeParamMap(ParamMap &pm)
{
    IncidentSource::MakeParamMap(pm);
    pm["SizeT"] = pint(&SizeT, 4096);
    pm["length"] = pdouble(&length, 1.);
    pm["width"] = pdouble(&width, 1.);
    pm["k0"] = pdouble(&k0, 2 * 3.14159265358979323846);
    pm["Shift"] = pdouble(&Shift, 0);
    pm["Time"] = pdouble(&Time, 0);
    pm["amp"] = pdouble(&amp, 1);
    pm["eps"] = pdouble(&eps, 1);
    pm["toff"] = pint(&toff, 0);
}

mpx_factory GaussInject::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_GAUSS_INJECT);
    f.d = distance;
    f.gsize_t = SizeT;
    f.length = length;
    f.width = width;
    f.k0 = k0;
    f.gshift = Shift;
    f.gtime = Time;
    f.amp = amp;
    f.eps = eps;
    f.toff = toff;
    return f;
}

void FocusedPulseInject::MakeParamMap(ParamMap &pm)
{
    IncidentSource::MakeParamMap(pm);
    pm["length"] = pdouble(&length, 1.);
    pm["width"] = pdouble(&width, 1.);
    pm["om0"] = pdouble(&om0, 2 * 3.14159265358979323846);
    pm["TShift"] = pdouble(&TShift, 0);
    pm["ZShift"] = pdouble(&ZShift, 0);
    pm["Phase"] = pdouble(&Phase, 0);
    pm["OversampleX"] = pint(&OversampleX, 1);
    pm["OversampleY"] = pint(&OversampleY, 1);
    pm["amp"] = pdouble(&amp, 1);
    pm["eps"] = pdouble(&eps, 1);
}

mpx_factory FocusedPulseInject::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_FOCUSED_PULSE);
    f.d = distance;
    f.length = length;
    f.width = width;
    f.om0 = om0;
    f.tshift = TShift;
    f.zshift = ZShift;
    f.phase = Phase;
    f.amp = amp;
    f.eps = eps;
    return f;
}

void PlaneWaveSource::MakeParamMap(ParamMap &pm)
{
    IncidentSource::MakeParamMap(pm);
    pm["kx"] = pdouble(&kx, 0);
    pm["ky"] = pdouble(&ky, 0);
    pm["kz"] = pdouble(&kz, 1);
    pm["Hx"] = pdouble(&Hx, 0);
    pm["Hy"] = pdouble(&Hy, 0);
    pm["Hz"] = pdouble(&Hz, 0);
    pm["Hx_bg"] = pdouble(&Hx_bg, 0);
    pm["Hy_bg"] = pdouble(&Hy_bg, 0);
    pm["Hz_bg"] = pdouble(&Hz_bg, 0);
    pm["Ex_bg"] = pdouble(&Ex_bg, 0);
    pm["Ey_bg"] = pdouble(&Ey_bg, 0);
    pm["Ez_bg"] = pdouble(&Ez_bg, 0);
    pm["ramp"] = pdouble(&ramp, 0.5);
    pm["eps"] = pdouble(&eps, 1);
}

mpx_factory PlaneWaveSource::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_PLANE_WAVE);
    f.d = distance;
    f.k[0] = kx;
    f.k[1] = ky;
    f.k[2] = kz;
    f.h[0] = Hx;
    f.h[1] = Hy;
    f.h[2] = Hz;
    f.h_bg[0] = Hx_bg;
    f.h_bg[1] = Hy_bg;
    f.h_bg[2] = Hz_bg;
    f.e_bg[0] = Ex_bg;
    f.e_bg[1] = Ey_bg;
    f.e_bg[2] = Ez_bg;
    f.ramp = ramp;
    f.eps = eps;
    return f;
}

void PlaneGaussSource::MakeParamMap(ParamMap &pm)
{
    IncidentSource::MakeParamMap(pm);
    pm["kx"] = pdouble(&kx, 0);
    pm["ky"] = pdouble(&ky, 0);
    pm["kz"] = pdouble(&kz, 1);
    pm["Hx"] = pdouble(&Hx, 0);
    pm["Hy"] = pdouble(&Hy, 0);
    pm["Hz"] = pdouble(&Hz, 0);
    pm["width"] = pdouble(&width, 10);
    pm["offset"] = pdouble(&offset, 40);
    pm["eps"] = pdouble(&eps, 1);
}

mpx_factory PlaneGaussSource::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_PLANE_GAUSS);
    f.d = distance;
    f.k[0] = kx;
    f.k[1] = ky;
    f.k[2] = kz;
    f.h[0] = Hx;
    f.h[1] = Hy;
    f.h[2] = Hz;
    f.gwidth = width;
    f.goffset = offset;
    f.eps = eps;
    return f;
}

// A factory kind the GPU solver does not provide: building it is a setup error.
template <const char *Name>
struct Unsupported : CurrentFactory {
    Unsupported() { throw SetupError(std::string(Name) + " is not provided by the GPU solver"); }
    mpx_factory describe() const override { return mpx_factory(); }

  protected:
    void MakeParamMap(ParamMap &) override {}
};
static const char kSideInject[] = "side_inject";

static void push_factory(void *owner, Rebuildable *r)
{
    static_cast<FDTD_PLRCSolverHIP *>(owner)->currentFactories.push_back(static_cast<CurrentFactory *>(r));
}

static void set_plasma(void *owner, Rebuildable *r)
{
    FDTD_PLRCSolverHIP *s = static_cast<FDTD_PLRCSolverHIP *>(owner);
    if (s->plasma) {
        delete r;
        throw SetupError("one plasma block per solver is provided by the GPU solver");
    }
    s->plasma = static_cast<PlasmaDensityBlock *>(r);
}

void PlasmaCurrentFactory::MakeParamMap(ParamMap &pm)
{
    pm["em"] = pdouble(&em, 1.0);
    pm["gamma"] = pdouble(&gamma, 0.1);
    pm["rho_bg"] = pdouble(&rho_bg, 0.0);
}

mpx_factory PlasmaCurrentFactory::describe() const
{
    mpx_factory f;
    mpx_factory_defaults(&f, MPX_FACTORY_PLASMA_CURRENT);
    f.em = em;
    f.gamma = gamma;
    f.rho_bg = rho_bg;
    return f;
}

void PlasmaDensityBlock::MakeParamMap(ParamMap &pm)
{
    pm["nu"] = pdouble(&nu, 1.0);
    pm["mpa"] = pdouble(&mpa, 1.0);
    pm["Wion"] = pdouble(&Wion, 1.0);
    pm["K"] = pint(&K, 5);
}

void FDTD_PLRCSolverHIP::MakeParamMap(ParamMap &pm)
{
    if (nonlin_) pm["chi"] = pdouble(&chi, 0.1);  // src/fdtd_plrc.cpp:461-465
    pm["eps"] = pdouble(&eps, 1.0);
    pm["L1eps"] = pdouble(&LEps[0], 0.0);
    pm["L2eps"] = pdouble(&LEps[1], 0.0);
    pm["L3eps"] = pdouble(&LEps[2], 0.0);
    pm["L1delta"] = pdouble(&LDelta[0], 0.0);
    pm["L2delta"] = pdouble(&LDelta[1], 0.0);
    pm["L3delta"] = pdouble(&LDelta[2], 0.0);
    pm["L1Om2"] = pdouble(&LOm2[0], 1.0);
    pm["L2Om2"] = pdouble(&LOm2[1], 1.0);
    pm["L3Om2"] = pdouble(&LOm2[2], 1.0);
    pm["cpml_border"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, CPMLBorder>(this, push_factory);
    pm["cpml_border_1d"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, CPMLBorderOneD>(this, push_factory);
    pm["short_pulse_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, ShortPulseInject>(this, push_factory);
    pm["wave_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, PlaneWaveSource>(this, push_factory);
    pm["plane_gauss_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, PlaneGaussSource>(this, push_factory);
    pm["plasma"] = pobject<FDTD_PLRCSolverHIP, Rebuildable, PlasmaDensityBlock>(this, set_plasma);
    pm["plasma_current"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, PlasmaCurrentFactory>(this, push_factory);
    pm["side_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, Unsupported<kSideInject>>(this, push_factory);
    pm["gauss_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, GaussInject>(this, push_factory);
    pm["focused_pulse_inject"] = pobject<FDTD_PLRCSolverHIP, CurrentFactory, FocusedPulseInject>(this, push_factory);
}

FDTD_PLRCSolverHIP::~FDTD_PLRCSolverHIP()
{
    for (CurrentFactory *f : currentFactories) delete f;
    delete plasma;
}

void FDTD_PLRCSolverHIP::describe(mpx_desc &d) const
{
    d.solver = nonlin_ ? MPX_SOLVER_PLRC_NONLIN : MPX_SOLVER_PLRC_LIN;
    d.eps = eps;
    d.chi = nonlin_ ? chi : 0.0;
    for (int n = 0; n < 3; ++n) {
        d.l_eps[n] = LEps[n];
        d.l_delta[n] = LDelta[n];
        d.l_om2[n] = LOm2[n];
    }
    d.n_factories = 0;
    for (CurrentFactory *f : currentFactories) {
        if (d.n_factories >= MPX_MAX_FACTORIES) throw SetupError("too many current factories");
        d.factories[d.n_factories++] = f->describe();
    }
    d.plasma = plasma ? 1 : 0;
    if (plasma) {
        d.plasma_k = plasma->K;
        d.plasma_nu = plasma->nu;
        d.plasma_mpa = plasma->mpa;
        d.plasma_wion = plasma->Wion;
    }
}

void FDTD_SimpleHIP::MakeParamMap(ParamMap &pm)
{
    pm["eps"] = pdouble(&eps, 1.0);
    if (solver_kind_ != MPX_SOLVER_PLAIN) pm["chi"] = pdouble(&chi, 0.1);
    if (solver_kind_ == MPX_SOLVER_DISP) {
        pm["L1eps2"] = pdouble(&LEps2[0], 0.0);
        pm["L2eps2"] = pdouble(&LEps2[1], 0.0);
        pm["L3eps2"] = pdouble(&LEps2[2], 0.0);
        pm["L1Om2"] = pdouble(&LOm2[0], 0.0);
        pm["L2Om2"] = pdouble(&LOm2[1], 0.0);
        pm["L3Om2"] = pdouble(&LOm2[2], 0.0);
    }
}

void FDTD_SimpleHIP::describe(mpx_desc &d) const
{
    d.solver = solver_kind_;
    d.eps = eps;
    d.chi = chi;
    for (int n = 0; n < 3; ++n) {
        d.d_eps2[n] = LEps2[n];
        d.d_om2[n] = LOm2[n];
    }
    d.n_factories = 0;
}

// ---------------------------------------------------------------------------------------
// FieldSolver
// ---------------------------------------------------------------------------------------

FieldSolver::~FieldSolver()
{
    if (solver_) mpx_destroy(solver_);
}

void FieldSolver::addCurrent(Current *)
{
    // src/fieldsolver.h:15-23
    throw SetupError("Additional currents cannot be used when using this Field Solver");
}

void FieldSolver::addMagCurrent(Current *)
{
    throw SetupError("Additional magnetic currents cannot be used when using this Field Solver");
}

static mpx_desc full_desc(const FieldSolver &s, const Globals &g, const Boundary *b)
{
    mpx_desc d;
    mpx_desc_defaults(&d);
    d.nx = g.GridX;
    d.ny = g.GridY;
    d.nz = g.GridZ;
    d.dx = g.GridDX;
    d.dy = g.GridDY;
    d.dz = g.GridDZ;
    d.dt = g.DT;
    d.boundary = b ? b->kind() : MPX_BOUNDARY_PERIODIC;
    d.precision = s.precision;
    d.path = s.path;
    s.describe(d);
    return d;
}

void FieldSolver::initStorage(Storage *)
{
    if (!globals) throw SetupError("No globals specified! Must exit!");
    mpx_desc d = full_desc(*this, *globals, boundary);
    if (solver_) {
        mpx_destroy(solver_);
        solver_ = nullptr;
    }
    if (mpx_create(&d, &solver_) != MPX_OK) throw SetupError(mpx_last_error());
}

void FieldSolver::stepSchemeInit(double)
{
    if (mpx_step_init(solver_) != MPX_OK) throw SetupError(mpx_last_error());
}

void FieldSolver::stepScheme(double)
{
    // the sources read the driver's time index (Process::time, src/incsource.t:46)
    if (time && mpx_set_time(solver_, *time) != MPX_OK) throw SetupError(mpx_last_error());
    if (mpx_step(solver_, 1) != MPX_OK) throw SetupError(mpx_last_error());
}

// ---------------------------------------------------------------------------------------
// Storage (host view)
// ---------------------------------------------------------------------------------------

std::vector<double> Storage::getGrid(const std::string &name) const
{
    const mpx_desc *unused = nullptr;
    (void)unused;
    Globals *g = solver->globals;
    std::vector<double> v((size_t)(g->GridX + 2) * (g->GridY + 2) * (g->GridZ + 2));
    if (mpx_get_field(solver->handle(), name.c_str(), v.data(), v.size()) != MPX_OK) throw SetupError(mpx_last_error());
    return v;
}

std::vector<double> Storage::getBorderLayer(const std::string &name, int dir, int lo[3], int hi[3]) const
{
    if (mpx_get_layer(solver->handle(), name.c_str(), dir, nullptr, 0, lo, hi) != MPX_OK)
        throw SetupError(mpx_last_error());
    std::vector<double> v((size_t)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1));
    if (mpx_get_layer(solver->handle(), name.c_str(), dir, v.data(), v.size(), lo, hi) != MPX_OK)
        throw SetupError(mpx_last_error());
    return v;
}

double Storage::energy() const
{
    double e = 0;
    if (mpx_energy(solver->handle(), &e) != MPX_OK) throw SetupError(mpx_last_error());
    return solver->boundary ? solver->boundary->SumReduce(e) : e;
}

// ---------------------------------------------------------------------------------------
// FieldSimulation / Process
// ---------------------------------------------------------------------------------------

static void add_fielddiag(void *owner, Rebuildable *r)
{
    static_cast<FieldSimulation *>(owner)->fieldDiags.push_back(static_cast<FieldDiag *>(r));
}

static void add_otherdiag(void *owner, Rebuildable *r)
{
    static_cast<FieldSimulation *>(owner)->otherDiags.push_back(static_cast<SimpleFieldDiag *>(r));
}

static void set_solver(void *owner, Rebuildable *r)
{
    auto *fs = static_cast<FieldSimulation *>(owner);
    delete fs->solver;
    fs->solver = static_cast<FieldSolver *>(r);
}
static void set_boundary(void *owner, Rebuildable *r)
{
    auto *fs = static_cast<FieldSimulation *>(owner);
    delete fs->boundary;
    fs->boundary = static_cast<Boundary *>(r);
}

template <bool NL>
static Rebuildable *make_plrc()
{
    return new FDTD_PLRCSolverHIP(NL);
}
template <int K>
static Rebuildable *make_simple()
{
    return new FDTD_SimpleHIP(K);
}

void FieldSimulation::MakeParamMap(ParamMap &pm)
{
    // solver, boundary registries of src/fieldsim.cpp:177-214
    auto obj = [&](Rebuildable *(*mk)(), void (*sink)(void *, Rebuildable *)) {
        Param q;
        q.kind = Param::OBJECT;
        q.make = mk;
        q.sink = sink;
        q.owner = this;
        return q;
    };
    pm["fdtd-plain"] = obj(&make_simple<MPX_SOLVER_PLAIN>, set_solver);
    pm["fdtd-nl"] = obj(&make_simple<MPX_SOLVER_NL>, set_solver);
    pm["fdtd-disp"] = obj(&make_simple<MPX_SOLVER_DISP>, set_solver);
    pm["fdtd-plrc-lin"] = obj(&make_plrc<false>, set_solver);
    pm["fdtd-plrc-nonlin"] = obj(&make_plrc<true>, set_solver);
    pm["single_periodic"] = obj(&make_new<SinglePeriodicBoundary>, set_boundary);
    pm["single_xy_periodic"] = obj(&make_new<SingleXYPeriodicBoundary>, set_boundary);
    // the MPI boundaries wrap the same periodic domain; on one GPU the whole domain is local
    pm["mpi_slice_periodic"] = obj(&make_new<SinglePeriodicBoundary>, set_boundary);
    pm["mpi_block_periodic"] = obj(&make_new<SinglePeriodicBoundary>, set_boundary);
    // diagnostics registry, src/fieldsim.cpp:169-171
    pm["fielddiag"] = obj(&make_new<FieldDiag>, add_fielddiag);
    pm["slicediag"] = obj(&make_new<FieldSliceDiag>, add_otherdiag);
    pm["linediag"] = obj(&make_new<FieldLineDiag>, add_otherdiag);
    pm["energy"] = obj(&make_new<FieldEnergyDiag>, add_otherdiag);
    pm["frequency"] = obj(&make_new<FrequencyDiag>, add_otherdiag);
    pm["followmax"] = obj(&make_new<FollowMax>, add_otherdiag);
}

FieldSimulation::~FieldSimulation()
{
    for (FieldDiag *d : fieldDiags) delete d;  // each finishes its writes first
    for (SimpleFieldDiag *d : otherDiags) delete d;
    delete solver;
    delete boundary;
}

void FieldSimulation::executeDiagnostics()
{
    // DiagnosticManager::execute (src/process.cpp:29); each diagnostic writes its own files,
    // so the order among them does not show in the outputs
    for (FieldDiag *d : fieldDiags) d->execute();
    for (SimpleFieldDiag *d : otherDiags) d->execute();
}

void FieldSimulation::finishDiagnostics()
{
    for (FieldDiag *d : fieldDiags) {
        d->finish();
        if (!d->error.empty()) throw SetupError(d->error);
    }
    for (SimpleFieldDiag *d : otherDiags) {
        d->finish();
        if (!d->error.empty()) throw SetupError(d->error);
    }
}

// ---------------------------------------------------------------------------------------
// text and slice diagnostics
// ---------------------------------------------------------------------------------------

SimpleFieldDiag::~SimpleFieldDiag() { close(); }

void SimpleFieldDiag::MakeParamMap(ParamMap &pm)
{
    // DiagnosticInterface::MakeParamMap, src/diagnostic.cpp:34-47
    pm["file"] = pstring(&fname, "");
    pm["append"] = pstring(&append, "n");
    pm["interval"] = pint(&interval, 100);
}

void SimpleFieldDiag::init(FieldSolver *solver, Globals *g, int *time_)
{
    gpu = solver->handle();
    globals = g;
    time = time_;
    if (interval < 1) throw SetupError("Diagnostic: interval must be positive");
}

std::string SimpleFieldDiag::parsedFileName() const
{
    // DiagnosticInterface::parsedFileName, src/diagnostic.cpp:53-72 (one process: id 0)
    std::string parsed = fname;
    size_t pos = parsed.find("#p");
    if (pos != std::string::npos) parsed.replace(pos, 2, "0");
    pos = parsed.find("#t");
    if (pos != std::string::npos) parsed.replace(pos, 2, std::to_string(time ? *time : 0));
    return parsed;
}

bool SimpleFieldDiag::open(const std::string &path)
{
    close();
    out = new std::ofstream(path.c_str());
    if (!*out) {
        // the reference's std::ofstream stays in its failed state and drops the writes (an
        // empty "file", the default, gives no output); warn once instead of failing the run
        if (!warned) std::fprintf(stderr, "Diagnostic: cannot open '%s'; its output is dropped\n", path.c_str());
        warned = true;
        close();
        return false;
    }
    return true;
}

void SimpleFieldDiag::close()
{
    if (out) {
        out->close();
        delete out;
        out = nullptr;
    }
}

void SimpleFieldDiag::finish()
{
    if (out) out->flush();
}

std::vector<double> SimpleFieldDiag::grid(const std::string &name)
{
    std::vector<double> v((size_t)(globals->GridX + 2) * (globals->GridY + 2) * (globals->GridZ + 2));
    if (mpx_get_field(gpu, name.c_str(), v.data(), v.size()) != MPX_OK && error.empty())
        error = std::string("Diagnostic: ") + mpx_last_error();
    return v;
}

void SimpleFieldDiag::execute()
{
    // DiagnosticInterface::execute, src/diagnostic.cpp:12-31
    if (!gpu) return;
    calculate();
    const bool appending = !append.empty() && append[0] == 'y';
    if (t == 0 && appending && active()) open(parsedFileName());
    if (t % interval == 0 && active()) {
        bool ok = appending || open(parsedFileName());
        if (ok) write();
        if (!appending) close();
    }
    ++t;
}

void FieldEnergyDiag::write()
{
    // FieldEnergyDiag::calculate / write, src/fielddiag.cpp:244-286
    double energy = 0;
    if (mpx_energy(gpu, &energy) != MPX_OK && error.empty()) error = std::string("Diagnostic: ") + mpx_last_error();
    if (out) *out << *time << " " << energy << std::endl;
}

void FieldLineDiag::MakeParamMap(ParamMap &pm)
{
    // FieldLineDiag::MakeParamMap, src/fielddiag.cpp:141-153
    SimpleFieldDiag::MakeParamMap(pm);
    pm["field"] = pstring(&fieldId, "");
    pm["dim"] = pstring(&direction, "x");
    pm["posx"] = pint(&posx, 5);
    pm["posy"] = pint(&posy, 5);
    pm["posz"] = pint(&posz, 5);
    pm["format"] = pstring(&format, "lines");
}

static int line_dim(const std::string &d) { return d.empty() ? 0 : (d[0] == 'y' ? 1 : (d[0] == 'z' ? 2 : 0)); }

bool FieldLineDiag::active() const
{
    // FieldLineDiag::fetchField, src/fielddiag.cpp:155-192: the line lies strictly inside the
    // storage box in both transverse dimensions
    const int dim = line_dim(direction), t1 = dim == 0 ? 1 : 0, t2 = dim == 2 ? 1 : 2;
    const int pos[3] = {posx, posy, posz}, high[3] = {globals->GridX + 1, globals->GridY + 1, globals->GridZ + 1};
    return pos[t1] > 0 && pos[t1] < high[t1] && pos[t2] > 0 && pos[t2] < high[t2];
}

bool FieldLineDiag::open(const std::string &path)
{
    if (!SimpleFieldDiag::open(path)) return false;
    out->precision(10);
    return true;
}

void FieldLineDiag::write()
{
    // FieldLineDiag::write, src/fielddiag.cpp:194-217
    if (!out) return;
    const int dim = line_dim(direction);
    const int ny2 = globals->GridY + 2, nz2 = globals->GridZ + 2, high = (dim == 0 ? globals->GridX : dim == 1 ? globals->GridY : globals->GridZ) + 1;
    std::vector<double> f = grid(fieldId);
    int i[3] = {posx, posy, posz};
    if (format != "gnu") {
        std::string sep;
        for (i[dim] = 0; i[dim] <= high; ++i[dim]) {
            *out << sep << f[((size_t)i[0] * ny2 + i[1]) * nz2 + i[2]];
            sep = " ";
        }
        *out << std::endl;
    } else {
        for (i[dim] = 0; i[dim] <= high; ++i[dim]) *out << i[dim] << ' ' << f[((size_t)i[0] * ny2 + i[1]) * nz2 + i[2]] << std::endl;
        *out << std::endl;
    }
}

FieldSliceDiag::~FieldSliceDiag() { close(); }

void FieldSliceDiag::MakeParamMap(ParamMap &pm)
{
    // FieldSliceDiag::MakeParamMap, src/fielddiag.cpp:45-53
    SimpleFieldDiag::MakeParamMap(pm);
    pm["field"] = pstring(&fieldId, "");
    pm["plane"] = pstring(&plane, "x");
    pm["pos"] = pint(&pos, 5);
}

bool FieldSliceDiag::active() const
{
    // FieldSliceDiag::fetchField, src/fielddiag.cpp:55-93: the plane lies strictly inside
    const int normal = line_dim(plane);
    const int high = (normal == 0 ? globals->GridX : normal == 1 ? globals->GridY : globals->GridZ) + 1;
    return pos > 0 && pos < high;
}

// ---------------------------------------------------------------------------------------
// diagnostics that read the grid after every step
// ---------------------------------------------------------------------------------------

SampledFieldDiag::~SampledFieldDiag()
{
    if (sampler) mpx_sampler_destroy(sampler);
}

void SampledFieldDiag::calculate()
{
    if (noGrid || !wanted()) return;
    if (!sampler) {
        int pos[3];
        samplerPos(pos);
        const int dim = samplerDim();
        const int n[3] = {globals->GridX, globals->GridY, globals->GridZ};
        len = dim < 0 ? 1 : n[dim] + 2;
        capacity = std::max(1, std::min(4096, (1 << 20) / len));
        const int rc = mpx_sampler_create(gpu, samplerField().c_str(), dim, pos, capacity, &sampler);
        if (rc == MPX_ERR_NAME) {  // storage->hasGrid(field) is false
            noGrid = true;
            return;
        }
        if (rc) {
            if (error.empty()) error = std::string("Diagnostic: ") + mpx_last_error();
            noGrid = true;
            return;
        }
        buf.resize((size_t)capacity * len);
    }
    if ((int)times.size() == capacity) catchUp();
    if (mpx_sampler_record(sampler) != MPX_OK) {
        if (error.empty()) error = std::string("Diagnostic: ") + mpx_last_error();
        return;
    }
    times.push_back(*time);
}

void SampledFieldDiag::catchUp()
{
    if (!sampler || times.empty()) return;
    int n = 0;
    if (mpx_sampler_drain(sampler, buf.data(), buf.size(), &n) != MPX_OK || n != (int)times.size()) {
        if (error.empty()) error = std::string("Diagnostic: ") + mpx_last_error();
        times.clear();
        return;
    }
    for (int r = 0; r < n; ++r) replay(times[r], buf.data() + (size_t)r * len);
    times.clear();
}

void FrequencyDiag::MakeParamMap(ParamMap &pm)
{
    // FrequencyDiag::MakeParamMap, src/freqdiag.cpp:72-95
    SimpleFieldDiag::MakeParamMap(pm);
    pm["field"] = pstring(&field, "Ex");
    pm["x"] = pint(&x, 1);
    pm["y"] = pint(&y, 1);
    pm["z"] = pint(&z, 1);
}

void FrequencyDiag::replay(int tstep, const double *values)
{
    // FrequencyDiag::calculate, src/freqdiag.cpp:45-70
    const double val = 0.15 * values[0];
    if (val * lastval <= 0 && std::fabs(val) * std::fabs(lastval) > 0) {
        const double d = lastval / (lastval - val);
        const double t = (tstep - 1 + d) * globals->DT;
        if (count < 0) {
            firstzero = t;
            count = 0;
        } else {
            ++count;
            frequency = (0.5 * count) / (t - firstzero);
        }
    }
    lastval = val;
}

void FrequencyDiag::write()
{
    // FrequencyDiag::write, src/freqdiag.cpp:20-29
    catchUp();
    if (count > lastcount) {
        lastcount = count;
        if (out) *out << *time << " " << frequency << std::endl;
    }
}

void FollowMax::MakeParamMap(ParamMap &pm)
{
    // FollowMax::MakeParamMap, src/followmax.cpp:115-125
    SimpleFieldDiag::MakeParamMap(pm);
    pm["field"] = pstring(&field, "Ex");
    pm["dim"] = pstring(&direction, "x");
    pm["posx"] = pint(&posx, 5);
    pm["posy"] = pint(&posy, 5);
    pm["posz"] = pint(&posz, 5);
    pm["start"] = pint(&start, 0);
}

int FollowMax::samplerDim() const { return line_dim(direction); }

bool FollowMax::open(const std::string &path)
{
    if (!SimpleFieldDiag::open(path)) return false;
    out->precision(10);
    return true;
}

bool FollowMax::wanted() const { return *time >= start; }

void FollowMax::setup()
{
    // FollowMax::init, src/followmax.cpp:32-71: the walk starts at pos[dim]
    if (ready) return;
    const int pos[3] = {posx, posy, posz};
    iposition = pos[samplerDim()];
    position = iposition;
    ready = true;
}

void FollowMax::replay(int, const double *line)
{
    // FollowMax::calculate, src/followmax.cpp:73-113, on the storage line low = 0 .. high = N + 1
    setup();
    const int high = samplerDim() == 0 ? globals->GridX + 1 : (samplerDim() == 1 ? globals->GridY + 1 : globals->GridZ + 1);
    const int low = 0;
    double mid, left, right;
    bool maxfound;
    do {
        int pm = iposition - 1, pp = iposition + 1;
        if (pm <= low) pm = high - 1;
        if (pp >= high) pp = low + 1;
        mid = line[iposition];
        left = line[pm];
        right = line[pp];
        if (right > mid) {
            iposition = pp;
            maxfound = false;
        } else if (left > mid) {
            iposition = pm;
            maxfound = false;
        } else {
            maxfound = true;
        }
    } while (!maxfound);
    const double offset = (left - right) / (2 * (left - 2 * mid + right));
    position = iposition + offset;
}

void FollowMax::write()
{
    // FollowMax::write, src/followmax.cpp:20-25
    setup();
    catchUp();
    if (*time < start) return;
    if (out) *out << *time << " " << position << std::endl;
}

bool FieldSliceDiag::open(const std::string &path)
{
    close();
    h5 = new HDFostream;
    if (!h5->open(path, 256)) {
        if (error.empty()) error = "Diagnostic: " + h5->error();
        delete h5;
        h5 = nullptr;
        return false;
    }
    return true;
}

void FieldSliceDiag::close()
{
    if (h5) {
        h5->close();
        delete h5;
        h5 = nullptr;
    }
}

void FieldSliceDiag::write()
{
    // FieldSliceDiag::write, src/fielddiag.cpp:95-118: slice(d1, d2) over the storage box of
    // the two in-plane dimensions (d1 < d2), streamed with d1 fastest
    if (!h5) return;
    const int normal = line_dim(plane), d1 = normal == 0 ? 1 : 0, d2 = normal == 2 ? 1 : 2;
    const int n[3] = {globals->GridX + 2, globals->GridY + 2, globals->GridZ + 2};
    std::vector<double> f = grid(fieldId), s((size_t)n[d1] * n[d2]);
    int i[3];
    i[normal] = pos;
    for (i[d2] = 0; i[d2] < n[d2]; ++i[d2])
        for (i[d1] = 0; i[d1] < n[d1]; ++i[d1]) s[(size_t)i[d2] * n[d1] + i[d1]] = f[((size_t)i[0] * n[1] + i[1]) * n[2] + i[2]];
    if (!h5->write({(uint64_t)n[d2], (uint64_t)n[d1]}, s.data()) && error.empty()) error = "Diagnostic: " + h5->error();
}

// ---------------------------------------------------------------------------------------
// FieldDiag
// ---------------------------------------------------------------------------------------

struct FieldDiag::Impl {
    mpx_solver *gpu = nullptr;
    int *time = nullptr;
    uint64_t nx2 = 0, ny2 = 0, nz2 = 0;
    double *buf[2] = {nullptr, nullptr};  // page-locked, alternating between outputs
    int next = 0;
    std::thread writer;
    HDFostream out;  // the open file of an appending diagnostic
    std::mutex errorMutex;
};

void FieldDiag::fail(const std::string &m)
{
    std::lock_guard<std::mutex> lock(impl->errorMutex);
    if (error.empty()) error = m;
}

FieldDiag::~FieldDiag()
{
    finish();
    if (impl) {
        impl->out.close();
        for (double *b : impl->buf)
            if (b) mpx_host_free(b);
        delete impl;
    }
}

void FieldDiag::MakeParamMap(ParamMap &pm)
{
    // FieldDiag::MakeParamMap (src/fielddiag.cpp:9-15), DiagnosticInterface::MakeParamMap
    // (src/diagnostic.cpp:34-47)
    pm["field"] = pstring(&fieldId, "");
    pm["file"] = pstring(&fname, "");
    pm["append"] = pstring(&append, "n");
    pm["interval"] = pint(&interval, 100);
}

void FieldDiag::init(FieldSolver *solver, Globals *g, int *time)
{
    if (!impl) impl = new Impl;
    impl->gpu = solver->handle();
    impl->time = time;
    impl->nx2 = g->GridX + 2;
    impl->ny2 = g->GridY + 2;
    impl->nz2 = g->GridZ + 2;
    if (interval < 1) throw SetupError("FieldDiag: interval must be positive");
}

std::string FieldDiag::parsedFileName() const
{
    // DiagnosticInterface::parsedFileName, src/diagnostic.cpp:53-72 (one process: id 0)
    std::string parsed = fname;
    size_t pos = parsed.find("#p");
    if (pos != std::string::npos) parsed.replace(pos, 2, "0");
    pos = parsed.find("#t");
    if (pos != std::string::npos) parsed.replace(pos, 2, std::to_string(impl && impl->time ? *impl->time : 0));
    return parsed;
}

void FieldDiag::execute()
{
    // DiagnosticInterface::execute, src/diagnostic.cpp:12-31
    if (!impl || !impl->gpu) return;
    const bool appending = !append.empty() && append[0] == 'y';
    if (t == 0 && appending) {
        if (!impl->out.open(parsedFileName(), 256)) fail("FieldDiag: " + impl->out.error());
    }
    if (t % interval == 0) {
        const size_t count = impl->nx2 * impl->ny2 * impl->nz2;
        double *&b = impl->buf[impl->next];
        if (!b && mpx_host_alloc(count * sizeof(double), (void **)&b) != MPX_OK) {
            fail(std::string("FieldDiag: ") + mpx_last_error());
            ++t;
            return;
        }
        // queue the copy behind this step, then hand the previous output's write over
        mpx_fetch *h = nullptr;
        if (mpx_fetch_field(impl->gpu, fieldId.c_str(), b, count, &h) != MPX_OK) {
            fail(std::string("FieldDiag: ") + mpx_last_error());
            ++t;
            return;
        }
        if (impl->writer.joinable()) impl->writer.join();
        const std::string path = appending ? std::string() : parsedFileName();
        const std::vector<uint64_t> dims = {impl->nz2, impl->ny2, impl->nx2};
        const double *data = b;
        impl->writer = std::thread([this, h, path, dims, data, appending]() {
            if (mpx_fetch_wait(h) != MPX_OK) {
                fail(std::string("FieldDiag: ") + mpx_last_error());
                return;
            }
            if (appending) {
                if (impl->out.good() && !impl->out.write(dims, data))
                    fail("FieldDiag: " + impl->out.error());
                return;
            }
            HDFostream o;
            if (!o.open(path, 1) || !o.write(dims, data)) fail("FieldDiag: " + o.error());
            o.close();
        });
        impl->next ^= 1;
    }
    ++t;
}

void FieldDiag::finish()
{
    if (impl && impl->writer.joinable()) impl->writer.join();
}

void FieldSimulation::init(Globals *g, int *time)
{
    // src/fieldsim.cpp:31-107 (initialisers and diagnostics are not part of this path)
    if (!solver) throw SetupError("No field solver specified");
    if (!boundary) boundary = new SinglePeriodicBoundary();
    dt = g->DT;
    solver->globals = g;
    solver->boundary = boundary;
    solver->time = time;
    storage.reset(new Storage(solver));
    solver->initStorage(storage.get());
    for (FieldDiag *d : fieldDiags) d->init(solver, g, time);  // FieldDiag::fetchField
    for (SimpleFieldDiag *d : otherDiags) d->init(solver, g, time);
}

void FieldSimulation::execute() { solver->stepScheme(dt); }

static void set_globals(void *owner, Rebuildable *r)
{
    auto *p = static_cast<Process *>(owner);
    delete p->globals;
    p->globals = static_cast<Globals *>(r);
}
static void set_fieldsim(void *owner, Rebuildable *r)
{
    auto *p = static_cast<Process *>(owner);
    delete p->fieldSim;
    p->fieldSim = static_cast<FieldSimulation *>(r);
}

void Process::MakeParamMap(ParamMap &pm)
{
    Param g;
    g.kind = Param::OBJECT;
    g.make = &make_new<Globals>;
    g.sink = set_globals;
    g.owner = this;
    pm["globals"] = g;
    Param f = g;
    f.make = &make_new<FieldSimulation>;
    f.sink = set_fieldsim;
    pm["fieldsim"] = f;
}

Process::~Process()
{
    delete fieldSim;
    delete globals;
}

std::string Process::RebuildAll(std::istream &in)
{
    // src/process.cpp:45-59
    std::string t = Rebuild(in);
    if (!globals) throw SetupError("No globals specified! Must exit!");
    if (!fieldSim) throw SetupError("No field simulation specified! Must exit!");
    return t;
}

void Process::init() { fieldSim->init(globals, &time); }

void Process::run(int steps, bool verbose)
{
    // src/process.cpp:17-31
    int T = steps < 0 ? globals->TotalTime : time + steps;
    for (; time < T; ++time) {
        if (verbose) std::cout << "Cycle " << time << std::endl << std::flush;
        fieldSim->execute();
        fieldSim->executeDiagnostics();
    }
    fieldSim->finishDiagnostics();  // every file written so far is complete on return
}

}  // namespace mpulse_hip

// ---------------------------------------------------------------------------------------
// C entry points
// ---------------------------------------------------------------------------------------

using namespace mpulse_hip;

struct mph_sim {
    Process process;
};

static thread_local std::string g_mph_err;

static int mph_fail(const std::string &m)
{
    g_mph_err = m;
    return MPX_ERR_INVALID;
}

extern "C" {

const char *mph_last_error(void) { return g_mph_err.c_str(); }

int mph_parse_setup(const char *text, mpx_desc *out, int *total_steps)
{
    if (!text || !out) return mph_fail("null argument");
    try {
        Process p;
        std::istringstream in(text);
        p.RebuildAll(in);
        if (!p.fieldSim->solver) return mph_fail("No field solver specified");
        *out = full_desc(*p.fieldSim->solver, *p.globals, p.fieldSim->boundary);
        if (total_steps) *total_steps = p.globals->TotalTime;
        return MPX_OK;
    } catch (const SetupError &e) {
        return mph_fail(e.msg);
    }
}

int mph_create(const char *text, int precision, int path, mph_sim **out)
{
    if (!text || !out) return mph_fail("null argument");
    *out = nullptr;
    mph_sim *sim = new mph_sim();
    try {
        std::istringstream in(text);
        sim->process.RebuildAll(in);
        if (!sim->process.fieldSim->solver) throw SetupError("No field solver specified");
        sim->process.fieldSim->solver->precision = precision;
        sim->process.fieldSim->solver->path = path;
        sim->process.init();
    } catch (const SetupError &e) {
        delete sim;
        return mph_fail(e.msg);
    }
    *out = sim;
    return MPX_OK;
}

int mph_run(mph_sim *sim, int steps)
{
    if (!sim) return mph_fail("null simulation");
    try {
        sim->process.run(steps, false);
    } catch (const SetupError &e) {
        return mph_fail(e.msg);
    }
    return MPX_OK;
}

mpx_solver *mph_solver(mph_sim *sim) { return sim ? sim->process.fieldSim->solver->handle() : nullptr; }

int mph_destroy(mph_sim *sim)
{
    delete sim;
    return MPX_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------------------
// HDFostream entry points
// ---------------------------------------------------------------------------------------

struct mph_h5 {
    HDFostream out;
};

int mph_h5_open(const char *path, int capacity, mph_h5 **out)
{
    if (!path || !out) return mph_fail("null argument");
    mph_h5 *h = new mph_h5;
    if (!h->out.open(path, capacity)) {
        std::string e = h->out.error();
        delete h;
        return mph_fail(e);
    }
    *out = h;
    return MPX_OK;
}

int mph_h5_set_block_name(mph_h5 *h, const char *name)
{
    if (!h || !name) return mph_fail("null argument");
    h->out.setBlockName(name);
    return MPX_OK;
}

int mph_h5_set_mtime(mph_h5 *h, unsigned int mtime)
{
    if (!h) return mph_fail("null argument");
    h->out.setModificationTime(mtime);
    return MPX_OK;
}

int mph_h5_write(mph_h5 *h, int rank, const unsigned long long *dims, const double *data, const int *extent,
                 int n_extent)
{
    if (!h || !dims || (!data && rank > 0)) return mph_fail("null argument");
    if (rank < 1 || rank > 3) return mph_fail("dataset rank must be 1..3");
    std::vector<uint64_t> d(dims, dims + rank);
    std::vector<int32_t> e;
    if (extent && n_extent > 0) e.assign(extent, extent + n_extent);
    if (!h->out.write(d, data, e)) return mph_fail(h->out.error());
    return MPX_OK;
}

int mph_h5_close(mph_h5 *h)
{
    if (!h) return mph_fail("null argument");
    h->out.close();
    delete h;
    return MPX_OK;
}
