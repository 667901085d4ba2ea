// C++ host layer over the C ABI of include/mpulse_hip.h.
//
// It mirrors the reference's plugin interfaces for the time-stepping path -- FieldSolver
// (src/fieldsolver.h:10-37), CurrentFactory (src/currentfactory.h:10-15), IncidentSource
// (src/incsource.h:16-31), Boundary (src/boundary.h:32-85), the FieldSimulation / Process
// driver (src/fieldsim.cpp:31-112, src/process.cpp:12-31) and the setup-file grammar of
// Rebuildable (src/rebuild.cpp:15-116) -- with the same registry keys, parameter keys and
// defaults.  The solvers keep their state on the GPU; stepScheme calls mpx_step.
#pragma once

#include <fstream>
#include <iosfwd>
#include <iostream>
#include <list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mpulse_hip.h"

namespace mpulse_hip {

// ---------------------------------------------------------------------------------------
// setup-file parameters: "key value" pairs and "name { ... }" blocks, "//" line comments
// ---------------------------------------------------------------------------------------

class Rebuildable;

// Reads one value of a parameter, or builds one child object.
struct Param {
    enum Kind { INT, DOUBLE, STRING, OBJECT } kind;
    int *ip = nullptr;
    double *dp = nullptr;
    std::string *sp = nullptr;
    // OBJECT: factory making the child; the child is handed to `sink`
    Rebuildable *(*make)() = nullptr;
    void (*sink)(void *owner, Rebuildable *child) = nullptr;
    void *owner = nullptr;
};
using ParamMap = std::map<std::string, Param>;

class Rebuildable {
  public:
    virtual ~Rebuildable() = default;
    // Rebuildable::Rebuild semantics: read tokens until "}" or end of input; unknown keys are
    // reported and skipped (blocks included); returns the next token after the block.
    std::string Rebuild(std::istream &in);
    std::vector<std::string> warnings;
    // kept, and printed on std::cerr as the reference's parser prints them (src/rebuild.cpp:34-50)
    void warn(const std::string &m)
    {
        warnings.push_back(m);
        std::cerr << m << std::endl;
    }

  protected:
    virtual void MakeParamMap(ParamMap &pm) = 0;
    static Param pint(int *p, int def)
    {
        *p = def;
        Param q;
        q.kind = Param::INT;
        q.ip = p;
        return q;
    }
    static Param pdouble(double *p, double def)
    {
        *p = def;
        Param q;
        q.kind = Param::DOUBLE;
        q.dp = p;
        return q;
    }
    static Param pstring(std::string *p, const std::string &def)
    {
        *p = def;
        Param q;
        q.kind = Param::STRING;
        q.sp = p;
        return q;
    }
};

// ---------------------------------------------------------------------------------------
// globals, storage, boundary
// ---------------------------------------------------------------------------------------

class Globals : public Rebuildable {  // src/globals.cpp:21-48
  public:
    int GridX = 128, GridY = 128, GridZ = 128;
    double GridDX = 0.1, GridDY = 0.1, GridDZ = 0.1, DT = 1e-2;
    int TotalTime = 100000;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class Boundary : public Rebuildable {  // src/boundary.h:32-85 (single GPU: identity reductions)
  public:
    virtual int kind() const = 0;  // MPX_BOUNDARY_*
    double AvgReduce(double v) const { return v; }
    double SumReduce(double v) const { return v; }
    double MaxReduce(double v) const { return v; }
    bool master() const { return true; }
    int procnum() const { return 0; }
    int getUniqueId() const { return 0; }

  protected:
    void MakeParamMap(ParamMap &) override {}
};
class SinglePeriodicBoundary : public Boundary {
  public:
    int kind() const override { return MPX_BOUNDARY_PERIODIC; }
};
class SingleXYPeriodicBoundary : public Boundary {
  public:
    int kind() const override { return MPX_BOUNDARY_XY_PERIODIC; }
};

class FieldSolver;

// Host view of the state the solver keeps on the GPU (Storage, src/storage.h:15-116):
// grids are fetched by reference name, z fastest over 0..N+1.
class Storage {
  public:
    explicit Storage(FieldSolver *s) : solver(s) {}
    std::vector<double> getGrid(const std::string &name) const;
    std::vector<double> getBorderLayer(const std::string &name, int dir, int lo[3], int hi[3]) const;
    double energy() const;
    FieldSolver *solver;
};

// ---------------------------------------------------------------------------------------
// current factories: translated into device descriptors
// ---------------------------------------------------------------------------------------

class CurrentFactory : public Rebuildable {  // src/currentfactory.h:10-15
  public:
    // the descriptor the GPU solver builds this factory's currents from
    virtual mpx_factory describe() const = 0;
};

class CPMLBorder : public CurrentFactory {  // src/cpml_border.cpp:141-151
  public:
    int thickness = 8;
    double kappaMax = 15, aMax = 0.25, sigmaMax = 3;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    virtual int kind() const { return MPX_FACTORY_CPML; }
};
class CPMLBorderOneD : public CPMLBorder {  // src/cpml_border.cpp:245-255
  protected:
    int kind() const override { return MPX_FACTORY_CPML_1D; }
};

class IncidentSource : public CurrentFactory {  // src/incsource.cpp:80-86
  public:
    int distance = 15;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class ShortPulseInject : public IncidentSource {  // src/shortpulseinject.cpp:35-50
  public:
    double length = 1, width = 1, om0 = 6.283185307179586, TShift = 0, ZShift = 0, Phase = 0, amp = 1, eps = 1;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

// focused_pulse_inject keys of src/focusedpulseinject.cpp:38-56; the field model is the
// library's own (see include/mpulse_hip.h, MPX_FACTORY_FOCUSED_PULSE).  The oversampling keys
// are accepted: the angular spectrum is summed on the domain's own periodic grid.
class FocusedPulseInject : public IncidentSource {
  public:
    double length = 1, width = 1, om0 = 6.283185307179586, TShift = 0, ZShift = 0, Phase = 0, amp = 1, eps = 1;
    int OversampleX = 1, OversampleY = 1;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class PlasmaCurrentFactory : public CurrentFactory {  // src/plasmacurrent.cpp:6-21
  public:
    double em = 1, gamma = 0.1, rho_bg = 0;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

// The "plasma" OptField of the PLRC solvers (PlasmaDensity, src/plasmadensity.cpp:62-73)
class PlasmaDensityBlock : public Rebuildable {
  public:
    double nu = 1, mpa = 1, Wion = 1;
    int K = 5;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class PlaneWaveSource : public IncidentSource {  // src/sources.cpp:58-82
  public:
    double kx = 0, ky = 0, kz = 1, Hx = 0, Hy = 0, Hz = 0, Hx_bg = 0, Hy_bg = 0, Hz_bg = 0;
    double Ex_bg = 0, Ey_bg = 0, Ez_bg = 0, ramp = 0.5, eps = 1;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class GaussInject : public IncidentSource {  // src/gaussinject.cpp:33-49
  public:
    int SizeT = 4096, toff = 0;
    double length = 1, width = 1, k0 = 6.283185307179586, Shift = 0, Time = 0, amp = 1, eps = 1;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class PlaneGaussSource : public IncidentSource {  // src/sources.cpp:225-242
  public:
    double kx = 0, ky = 0, kz = 1, Hx = 0, Hy = 0, Hz = 0, width = 10, offset = 40, eps = 1;
    mpx_factory describe() const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

// ---------------------------------------------------------------------------------------
// field solvers
// ---------------------------------------------------------------------------------------

class Current;  // host-computed currents are not accepted by the GPU solvers

class FieldSolver : public Rebuildable {  // src/fieldsolver.h:10-37
  public:
    ~FieldSolver() override;
    virtual void initStorage(Storage *storage);
    virtual void addCurrent(Current *);     // not supported: prints the reference's message, exits
    virtual void addMagCurrent(Current *);  // idem
    virtual void stepSchemeInit(double dt);
    virtual void stepScheme(double dt);

    // GPU execution choices (not setup-file keys of the reference)
    int precision = MPX_PREC_FP64;
    int path = MPX_PATH_FUSED;
    mpx_solver *handle() const { return solver_; }
    // fills the descriptor this solver hands to mpx_create
    virtual void describe(mpx_desc &d) const = 0;

    Globals *globals = nullptr;
    Boundary *boundary = nullptr;
    int *time = nullptr;  // Process::time

  protected:
    void MakeParamMap(ParamMap &) override {}
    mpx_solver *solver_ = nullptr;
};

// FDTD_PLRCSolver<FDTD_PLRC{Lin,Nonlin}Core> (src/fdtd_plrc.t:411-473, src/fdtd_plrc.cpp:461-465)
class FDTD_PLRCSolverHIP : public FieldSolver {
  public:
    explicit FDTD_PLRCSolverHIP(bool nonlin) : nonlin_(nonlin) {}
    double eps = 1, chi = 0.1;
    double LEps[3] = {0, 0, 0}, LDelta[3] = {0, 0, 0}, LOm2[3] = {1, 1, 1};
    std::list<CurrentFactory *> currentFactories;
    PlasmaDensityBlock *plasma = nullptr;
    ~FDTD_PLRCSolverHIP() override;
    void describe(mpx_desc &d) const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;

  private:
    bool nonlin_;
};

// FDTD_Plain / FDTD_Nonlinear / FDTD_Dispersion (src/fdtd_plain.h, src/fdtd_nl.cpp:146-153,
// src/fdtd_disp.cpp:252-272)
class FDTD_SimpleHIP : public FieldSolver {
  public:
    explicit FDTD_SimpleHIP(int solver) : solver_kind_(solver) {}
    double eps = 1, chi = 0.1, LEps2[3] = {0, 0, 0}, LOm2[3] = {0, 0, 0};
    void describe(mpx_desc &d) const override;

  protected:
    void MakeParamMap(ParamMap &pm) override;

  private:
    int solver_kind_;
};

// ---------------------------------------------------------------------------------------
// driver
// ---------------------------------------------------------------------------------------

// ---------------------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------------------

class HDFostream;

// FieldDiag (src/fielddiag.cpp:9-24) with the keys of DiagnosticInterface
// (src/diagnostic.cpp:34-68): "field", "file" (#p -> process id, #t -> time step), "append"
// (y/n), "interval".  execute() follows DiagnosticInterface::execute: output at its 0th call
// and every `interval` calls after; without append one file per output, with append one file
// whose datasets are data0, data1, ...  The grid goes out as the reference writes it: the
// storage box 0..N+1, x fastest, HDF5 dims (Nz+2, Ny+2, Nx+2), float64.  The copy is queued
// behind the step (mpx_fetch_field) into page-locked memory and a writer thread writes the
// file while the GPU steps on; finish() waits for the writes in flight.
class FieldDiag : public Rebuildable {
  public:
    ~FieldDiag() override;
    std::string fieldId, fname, append;
    int interval = 100;
    void init(FieldSolver *solver, Globals *g, int *time);
    void execute();
    void finish();
    std::string parsedFileName() const;  // DiagnosticInterface::parsedFileName
    std::string error;                   // first failure of a write, empty when all went out

  protected:
    void MakeParamMap(ParamMap &pm) override;

  private:
    struct Impl;
    Impl *impl = nullptr;
    int t = 0;
    void fail(const std::string &m);  // keeps the first failure (called from the writer thread too)
};

// Text and slice diagnostics with DiagnosticInterface's keys and rules (src/diagnostic.cpp:
// 12-72): "file" (#p, #t), "append", "interval"; output at the 0th call and every `interval`
// calls after; one file per output, or one file when appending.  The grids are read when an
// output is due (the reference computes every step and writes at the outputs: same values).
class SimpleFieldDiag : public Rebuildable {
  public:
    ~SimpleFieldDiag() override;
    std::string fname, append;
    int interval = 100;
    void init(FieldSolver *solver, Globals *g, int *time);
    void execute();
    void finish();
    std::string parsedFileName() const;
    std::string error;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    virtual bool active() const { return true; }
    virtual void calculate() {}                  // every step, before the output check
    virtual bool open(const std::string &path);  // text stream by default
    virtual void write() = 0;
    virtual void close();
    std::vector<double> grid(const std::string &name);  // storage box, z fastest
    mpx_solver *gpu = nullptr;
    Globals *globals = nullptr;
    int *time = nullptr;
    std::ofstream *out = nullptr;
    int t = 0;
    bool warned = false;
};

// FieldEnergyDiag (src/fielddiag.cpp:219-286): "time energy" lines, energy = sum over the
// interior of E^2 + B^2 times dx dy dz
class FieldEnergyDiag : public SimpleFieldDiag {
  protected:
    void write() override;
};

// FieldLineDiag (src/fielddiag.cpp:122-217): keys "field", "dim", "posx", "posy", "posz",
// "format" (lines | gnu); precision 10
class FieldLineDiag : public SimpleFieldDiag {
  public:
    std::string fieldId, direction, format;
    int posx = 5, posy = 5, posz = 5;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    bool active() const override;
    bool open(const std::string &path) override;
    void write() override;
};

// Base of the diagnostics that read the grid after every step (FrequencyDiag, FollowMax):
// the reads are queued on the GPU through an mpx_sampler and the reference's calculate()
// runs over them, in step order, when an output is due or the sampler is full.
class SampledFieldDiag : public SimpleFieldDiag {
  public:
    ~SampledFieldDiag() override;

  protected:
    void calculate() override;
    void catchUp();  // replays the recorded steps
    virtual bool wanted() const { return true; }                 // record at this step?
    virtual void replay(int time, const double *values) = 0;    // one step's calculate()
    virtual int samplerDim() const = 0;                         // -1: one cell
    virtual void samplerPos(int pos[3]) const = 0;
    virtual std::string samplerField() const = 0;
    bool noGrid = false;  // the storage has no such grid: calculate() does nothing

  private:
    mpx_sampler *sampler = nullptr;
    std::vector<int> times;  // step counter of each pending record
    std::vector<double> buf;
    int capacity = 0, len = 1;
};

// FrequencyDiag (src/freqdiag.cpp:5-95): keys "field" (Ex), "x", "y", "z" (1); the zero
// crossings of 0.15 G(x, y, z) between steps, linearly interpolated; a line "time frequency"
// at an output when a half period was counted since the last line
class FrequencyDiag : public SampledFieldDiag {
  public:
    std::string field;
    int x = 1, y = 1, z = 1;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    void write() override;
    void replay(int time, const double *values) override;
    int samplerDim() const override { return -1; }
    void samplerPos(int pos[3]) const override { pos[0] = x, pos[1] = y, pos[2] = z; }
    std::string samplerField() const override { return field; }

  private:
    double frequency = 0, lastval = 0, firstzero = 0;
    int count = -1, lastcount = 0;
};

// FollowMax (src/followmax.cpp:5-125): keys "field" (Ex), "dim" (x), "posx", "posy", "posz" (5),
// "start" (0); from step `start` on, walks along the line to a local maximum of the field and
// writes "time position" (position refined by a quadratic through the maximum and its two
// neighbours, the line's interior taken as periodic), precision 10
class FollowMax : public SampledFieldDiag {
  public:
    std::string field, direction;
    int posx = 5, posy = 5, posz = 5, start = 0;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    bool open(const std::string &path) override;
    void write() override;
    bool wanted() const override;
    void replay(int time, const double *values) override;
    int samplerDim() const override;
    void samplerPos(int pos[3]) const override { pos[0] = posx, pos[1] = posy, pos[2] = posz; }
    std::string samplerField() const override { return field; }

  private:
    void setup();        // FollowMax::init
    bool ready = false;
    int iposition = 0;
    double position = 0;
};

// FieldSliceDiag (src/fielddiag.cpp:26-118): keys "field", "plane" (x | y | z), "pos"; one
// float64 dataset per output through HDFostream, HDF5 dims (N2+2, N1+2) of the two in-plane
// dimensions d1 < d2, d1 fastest
class FieldSliceDiag : public SimpleFieldDiag {
  public:
    ~FieldSliceDiag() override;
    std::string fieldId, plane;
    int pos = 5;

  protected:
    void MakeParamMap(ParamMap &pm) override;
    bool active() const override;
    bool open(const std::string &path) override;
    void write() override;
    void close() override;
    HDFostream *h5 = nullptr;
};

class FieldSimulation : public Rebuildable {  // src/fieldsim.cpp:31-194
  public:
    ~FieldSimulation() override;
    FieldSolver *solver = nullptr;
    Boundary *boundary = nullptr;
    std::unique_ptr<Storage> storage;
    std::vector<FieldDiag *> fieldDiags;  // "fielddiag" blocks, in setup order
    std::vector<SimpleFieldDiag *> otherDiags;  // "slicediag", "linediag", "energy" blocks
    void init(Globals *g, int *time);
    void execute();
    void executeDiagnostics();  // DiagnosticManager::execute for the field diagnostics
    void finishDiagnostics();
    double dt = 0;

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

class Process : public Rebuildable {  // src/process.cpp:12-59
  public:
    ~Process() override;
    Globals *globals = nullptr;
    FieldSimulation *fieldSim = nullptr;
    int time = 0;
    std::string RebuildAll(std::istream &in);  // Process::Rebuild with its checks
    void init();
    void run(int steps = -1, bool verbose = true);  // steps < 0: T-total

  protected:
    void MakeParamMap(ParamMap &pm) override;
};

// Fatal-error convention of the reference (print, exit(-1)) as an exception the C entry
// points turn into a status code.
struct SetupError : std::exception {
    std::string msg;
    explicit SetupError(std::string m) : msg(std::move(m)) {}
    const char *what() const noexcept override { return msg.c_str(); }
};

}  // namespace mpulse_hip

extern "C" {
// Parse a setup text into the descriptor the GPU solver is created from (no GPU needed).
int mph_parse_setup(const char *text, mpx_desc *out, int *total_steps);
typedef struct mph_sim mph_sim;
// Parse, create the GPU solver (FieldSimulation::init) and run `steps` (T-total if < 0).
int mph_create(const char *text, int precision, int path, mph_sim **out);
int mph_run(mph_sim *sim, int steps);
mpx_solver *mph_solver(mph_sim *sim);
int mph_destroy(mph_sim *sim);
const char *mph_last_error(void);

// HDFostream (src/hdfstream.{h,t,cpp}) without the HDF5 library: open a file for up to
// `capacity` datasets, write float64 datasets of rank 1..3 (HDF5 dims slowest first, data in
// C order over them) named as HDFstream::getNextBlockName names them, optionally with the
// int32 "extent" attribute (n_extent > 0).  mtime 0: the time of each write.
typedef struct mph_h5 mph_h5;
int mph_h5_open(const char *path, int capacity, mph_h5 **out);
int mph_h5_set_block_name(mph_h5 *h, const char *name);
int mph_h5_set_mtime(mph_h5 *h, unsigned int mtime);
int mph_h5_write(mph_h5 *h, int rank, const unsigned long long *dims, const double *data, const int *extent,
                 int n_extent);
int mph_h5_close(mph_h5 *h);
}
