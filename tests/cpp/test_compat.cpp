// The VecGeom-shaped adapter (include/vgm/vecgeom_compat.h) used the way the
// reference's unit tests use UnplacedBox / VPlacedVolume and the navigator
// (test/unit_tests/TestBox.cpp:22-120 style assertions), with closed-form
// answers: a box 10 x 20 x 30 placed at the origin, and a world box holding
// one daughter box.  Exit status 0 on success.
#include <cmath>
#include <cstdio>
#include <vector>
#include "vgm/vecgeom_compat.h"

using namespace vgm::compat;

static int failures = 0;
#define EXPECT(c)                                                  \
  do {                                                             \
    if (!(c)) {                                                    \
      std::printf("FAILED line %d: %s\n", __LINE__, #c);           \
      ++failures;                                                  \
    }                                                              \
  } while (0)

int main()
{
  GeoManager geo(0);
  const int box = geo.Box(10., 20., 30.);
  PlacedVolume pv = geo.Place(geo.NewLogicalVolume(box));

  // points: origin, outside on +x, on the +y face, inside near -z
  std::vector<double> x{0, 20, 0, 0}, y{0, 0, 20, 0}, z{0, 0, 0, -25};
  std::vector<double> dx{1, -1, 0, 0}, dy{0, 0, -1, 0}, dz{0, 0, 0, -1};
  SOA3D<> P(x.data(), y.data(), z.data(), 4), D(dx.data(), dy.data(), dz.data(), 4);
  bool contains[4];
  Inside_t inside[4];
  double din[4], dout[4], sin_[4], sout[4];
  pv.Contains(P, contains);
  pv.Inside(P, inside);
  pv.DistanceToIn(P, D, nullptr, din);
  pv.DistanceToOut(P, D, nullptr, dout);
  pv.SafetyToIn(P, sin_);
  pv.SafetyToOut(P, sout);
  EXPECT(contains[0] && !contains[1] && contains[3]);
  EXPECT(inside[0] == kInside && inside[1] == kOutside && inside[2] == kSurface && inside[3] == kInside);
  EXPECT(dout[0] == 10.0);   // origin along +x to the x = 10 face
  EXPECT(din[1] == 10.0);    // (20,0,0) along -x to the x = 10 face
  EXPECT(dout[3] == 5.0);    // (0,0,-25) along -z to z = -30
  EXPECT(sout[0] == 10.0 && sin_[1] == 10.0 && sout[3] == 5.0);
  EXPECT(sin_[0] <= 0.0 && sout[1] <= 0.0);  // wrong-side safeties are not positive

  // navigation: world 100^3 holding a 10^3 box at (30, 0, 0)
  const int wshape = geo.Box(100., 100., 100.), dshape = geo.Box(10., 10., 10.);
  LogicalVolume wlv = geo.NewLogicalVolume(wshape), dlv = geo.NewLogicalVolume(dshape);
  PlacedVolume d = geo.PlaceDaughter(wlv, dlv, Transformation3D(30., 0., 0.));
  PlacedVolume world = geo.Place(wlv);
  Navigator nav(geo, world);
  EXPECT(nav.NumberOfStates() == 2);
  std::vector<double> gx{0, 30}, gy{0, 0}, gz{0, 0}, ux{1, 1}, uy{0, 0}, uz{0, 0};
  SOA3D<> G(gx.data(), gy.data(), gz.data(), 2), U(ux.data(), uy.data(), uz.data(), 2);
  uint32_t state[2], last[2] = {0, 0};
  uint8_t boundary[2] = {0, 0};
  double step[2], safety[2];
  nav.LocateGlobalPoints(G, state);
  EXPECT(state[0] == 1 && state[1] == 2);
  nav.ComputeSafeties(G, state, safety);
  EXPECT(safety[0] == 20.0 && safety[1] == 10.0);
  nav.ComputeStepsAndPropagatedStates(G, U, nullptr, state, last, boundary, step);
  EXPECT(step[0] == 20.0 && state[0] == 2 && boundary[0] == 1);  // enters the daughter
  EXPECT(step[1] == 10.0 && state[1] == 1 && boundary[1] == 1);  // leaves it into the world
  int dpv, level;
  uint32_t mother;
  nav.Decode(2, &dpv, &level, &mother);
  EXPECT(dpv == d.id() && level == 1 && mother == 1);
  // NavStatePath <-> index, and steps with safeties (the safety of the start
  // point first: the tracks are back at their start and located again)
  EXPECT(nav.PathToState({world.id(), d.id()}) == 2);
  EXPECT(nav.StateToPath(2) == std::vector<int>({world.id(), d.id()}));
  nav.LocateGlobalPoints(G, state);
  last[0] = last[1] = 0;
  boundary[0] = boundary[1] = 0;
  bool calc[2] = {true, false};
  nav.ComputeStepsAndSafetiesAndPropagatedStates(G, U, nullptr, state, last, boundary, step, calc, safety);
  EXPECT(safety[0] == 20.0 && safety[1] == 0.0);
  EXPECT(step[0] == 20.0 && state[0] == 2 && step[1] == 10.0 && state[1] == 1);

  // ComputeStep (no relocation): the first track's state is the daughter it
  // hits; the second leaves the daughter and keeps its state, on the boundary,
  // with the daughter as the last exited volume (PrepareOutState)
  nav.LocateGlobalPoints(G, state);
  last[0] = last[1] = 0;
  boundary[0] = boundary[1] = 0;
  nav.ComputeSteps(G, U, nullptr, state, last, boundary, step);
  EXPECT(step[0] == 20.0 && state[0] == 2 && boundary[0] == 1 && last[0] == 1);
  EXPECT(step[1] == 10.0 && state[1] == 2 && boundary[1] == 1 && last[1] == 2);
  // under a physics limit of 5 nothing is reached: states kept, off boundary
  double limit[2] = {5.0, 5.0};
  nav.LocateGlobalPoints(G, state);
  last[0] = last[1] = 0;
  nav.ComputeSteps(G, U, limit, state, last, boundary, step);
  EXPECT(step[0] == 5.0 && state[0] == 1 && boundary[0] == 0 && step[1] == 5.0 && state[1] == 2);
  // ComputeStepAndNextVolume is ComputeStepAndPropagatedState
  nav.LocateGlobalPoints(G, state);
  last[0] = last[1] = 0;
  boundary[0] = boundary[1] = 0;
  nav.ComputeStepsAndNextVolumes(G, U, nullptr, state, last, boundary, step);
  EXPECT(step[0] == 20.0 && state[0] == 2 && step[1] == 10.0 && state[1] == 1);

  // ApproachNextDaughter: from the origin along +x the daughter's box begins
  // at x = 20; the daughter has no daughters of its own, so its track keeps its step
  nav.LocateGlobalPoints(G, state);
  last[0] = last[1] = 0;
  double appr[2] = {1e300, 1e300};
  nav.ApproachNextDaughters(G, U, state, last, appr);
  EXPECT(appr[0] == 20.0 && appr[1] == 1e300);

  // a box minus a square bore along z
  const int outer = geo.Box(10., 10., 10.), bore = geo.Box(5., 5., 20.);
  PlacedVolume holed = geo.Place(geo.NewLogicalVolume(geo.Boolean(2, outer, nullptr, bore, nullptr)));
  std::vector<double> hx{0, 7}, hy{0, 0}, hz{0, 0};
  SOA3D<> H(hx.data(), hy.data(), hz.data(), 2);
  bool hin[2] = {true, false};
  holed.Contains(H, hin);
  EXPECT(!hin[0] && hin[1]);

  // a rod in a sleeve: touching along r = 5, so the shared cylinder is
  // inside the union; the outer wall's normal is radial
  const int rod = geo.Tube(0., 5., 10., 0., 2 * M_PI), sleeve = geo.Tube(5., 8., 10., 0., 2 * M_PI);
  PlacedVolume both = geo.Place(geo.NewLogicalVolume(geo.Boolean(0, rod, nullptr, sleeve, nullptr)));
  std::vector<double> qx{5, 8}, qy{0, 0}, qz{0, 0};
  SOA3D<> UP(qx.data(), qy.data(), qz.data(), 2);
  Inside_t uin[2] = {0, 0};
  both.Inside(UP, uin);
  EXPECT(uin[0] == 1 && uin[1] == 2);
  const double wall[3] = {0., 8., 0.};
  double nrm[3] = {0., 0., 0.};
  EXPECT(both.Normal(wall, nrm) && nrm[0] == 0. && nrm[1] == 1. && nrm[2] == 0.);

  // a hollow sphere: the shell's inner wall is 6 mm from the origin along +x
  PlacedVolume shell = geo.Place(geo.NewLogicalVolume(geo.Sphere(6., 8.)));
  std::vector<double> ox{0}, oy{0}, oz{0}, ex{1}, ey{0}, ez{0};
  SOA3D<> O(ox.data(), oy.data(), oz.data(), 1), E(ex.data(), ey.data(), ez.data(), 1);
  double din_shell = 0.;
  shell.DistanceToIn(O, E, nullptr, &din_shell);
  EXPECT(din_shell == 6.);

  // TestTorus2.cpp's t1 (tube radius 90 around a circle of radius 100),
  // entered along -x from (240, 0, 0): the outer surface at x = 190
  PlacedVolume ring = geo.Place(geo.NewLogicalVolume(geo.Torus(0., 90., 100.)));
  std::vector<double> tx{240}, ty{0}, tz{0}, mx{-1}, my{0}, mz{0};
  SOA3D<> TP(tx.data(), ty.data(), tz.data(), 1), TD(mx.data(), my.data(), mz.data(), 1);
  double din_ring = 0.;
  ring.DistanceToIn(TP, TD, nullptr, &din_ring);
  EXPECT(std::fabs(din_ring - 50.) < 1e-9);
  double n_ring[3];
  const double on_ring[3] = {190., 0., 0.};
  EXPECT(ring.Normal(on_ring, n_ring) && std::fabs(n_ring[0] - 1.) < 1e-12 && n_ring[1] == 0. && n_ring[2] == 0.);
  // a torus is not a boolean operand of this library
  bool torus_refused = false;
  try {
    geo.Boolean(2, geo.Box(300., 300., 100.), nullptr, geo.Torus(0., 90., 100.), nullptr);
  } catch (const Error &e) {
    torus_refused = e.code == VGM_EINVAL;
  }
  EXPECT(torus_refused);

  // errors surface as exceptions with the library's message
  bool threw = false;
  try {
    geo.Box(-1., 1., 1.);
  } catch (const Error &e) {
    threw = e.code == VGM_EINVAL;
  }
  EXPECT(threw);
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
  return failures ? 1 : 0;
}
