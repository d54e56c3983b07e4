// ref_harness.cpp -- C entry points around the reference's network class
// (network.cpp / network.h, compiled from /root/reference by ref.mk) so that
// tests can compare the oracle's fibrin forces, kick and graph construction
// with the reference's own code.  TEST INFRASTRUCTURE ONLY.
#include <vector>

#include "network.h"

extern "C" {

// Forces of network::getSpringForce (network.cpp:303-437) for links (a < b)
// with rest positions rest[3n] and current positions cur[3n].
int ref_spring_force(int n, const double *rest, const double *cur, int nl,
                     const int *links, double modulus, double mass,
                     double *force) {
  network net;
  net.linkr.assign(n, std::vector<int>());
  net.fibersize.assign(n, std::vector<double>());
  for (int i = 0; i < n; ++i) net.linkr[i].push_back(i);
  for (int l = 0; l < nl; ++l) {
    const int a = links[2 * l], b = links[2 * l + 1];
    net.linkr[a].push_back(b);
    net.fibersize[a].push_back(1.0);
    net.linkr[b].push_back(a);
    net.fibersize[b].push_back(1.0);
  }
  net.node.assign(rest, rest + 3 * n);
  std::vector<double> nodein(cur, cur + 3 * n);
  net.getSpringForce(nodein, 0, modulus, mass);
  for (int k = 0; k < 3 * n; ++k) force[k] = net.Force[k];
  return 0;
}

// network::solverkick (network.cpp:839-881) on vel[m] with force[m].
int ref_solverkick(int m, double *vel, const double *force, double dt,
                   double gamma) {
  network net;
  net.newp.assign(m, 0.0);
  net.newvel.assign(vel, vel + m);
  net.Force.assign(force, force + m);
  net.solverkick(dt, gamma);
  for (int k = 0; k < m; ++k) vel[k] = net.newvel[k];
  return 0;
}

// network::definelinks (network.cpp:197-218): counts[i] and, concatenated,
// the neighbour lists (including i itself) of every node.
int ref_definelinks(int n, const double *xyz, double dist_sq, int *counts,
                    int *lists, int cap) {
  network net;
  net.branchp.clear();
  for (int i = 0; i < n; ++i) net.putIntoVec(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  net.definelinks(dist_sq);
  int k = 0;
  for (int i = 0; i < n; ++i) {
    counts[i] = (int)net.linkr[i].size();
    for (int j : net.linkr[i]) {
      if (k >= cap) return -1;
      lists[k++] = j;
    }
  }
  return k;
}
}
