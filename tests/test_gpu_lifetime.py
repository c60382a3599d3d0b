"""Lifetime of the library's mesh object and of its scratch under graphs.

The mesh is shared by the objects built on it (leapfrogs, vertical
coordinates, graphs, meshes borrowing its stream); each holds a reference,
so any finalization order is safe.  Scratch that captured graphs use is never
freed under them: growth during a capture is refused, growth while graphs
exist keeps their buffers.  The Python mirror holds no reference cycles
through its meshes, so dropping the last reference frees a mesh at once."""
import gc
import weakref

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

A = 6.371e6


def test_objects_keep_their_mesh_alive_in_any_destroy_order(jgcm):
    N, nlat = 21, 32
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 3, A)
    assert mesh.lifetime() == (1, 0)
    I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, mesh.laplacian_eig, 0.5, 600, True, 0, 86400)
    vc = jgcm.Vert_Coordinate(2 * nlat, nlat, 3, "even_sigma", "simmons_and_burridge",
                              "second_centered_wts", 1e5, mesh=mesh)
    borrower = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 1, A)
    jgcm._lib.call("jgcm_mesh_share_stream", borrower.handle, mesh.handle)
    g = mesh.grid_zeros()
    s = mesh.spectral_zeros()
    graph = jgcm.Step_Graph(mesh, lambda: jgcm.Trans_Grid_To_Spherical(mesh, g, s), nfields=3)
    # creator + leapfrog + vertical coordinate + borrower + graph
    assert mesh.lifetime() == (5, 1)
    # the creator lets go first: every later destroy still finds its mesh
    handle = mesh.handle
    lib = jgcm._lib
    lib.call("jgcm_mesh_destroy", handle)
    mesh._h = None
    import ctypes
    r, lg = ctypes.c_int(), ctypes.c_int()
    lib.call("jgcm_mesh_lifetime", handle, ctypes.byref(r), ctypes.byref(lg))
    assert (r.value, lg.value) == (4, 1)
    graph.launch()                    # launches on the (still alive) stream
    graph.close()
    lib.call("jgcm_mesh_lifetime", handle, ctypes.byref(r), ctypes.byref(lg))
    assert (r.value, lg.value) == (3, 0)
    borrower.synchronize()            # uses the creator's stream
    borrower.close()
    del vc
    I.close()                         # the last reference: the mesh is freed here


def test_meshes_are_freed_without_the_cycle_collector(jgcm):
    """No reference cycle runs through a mesh: tables point back to it weakly,
    so dropping the last reference frees it at once."""
    N, nlat = 21, 32
    was = gc.isenabled()
    gc.disable()
    try:
        mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, 2, A)
        I = jgcm.Filtered_Leapfrog(0.04, 4, 1e-4, mesh.laplacian_eig, 0.5, 600, True, 0, 86400)
        assert I.laplacian_eigen.mesh is mesh
        ref = weakref.ref(mesh)
        del mesh
        assert ref() is not None       # the leapfrog still holds it
        del I
        assert ref() is None
    finally:
        if was:
            gc.enable()


def test_scratch_growth_is_refused_in_a_capture_and_kept_for_live_graphs(jgcm):
    N, nlat, nd = 21, 32, 2
    mesh = jgcm.Spectral_Spherical_Mesh(N, N + 1, nlat, 2 * nlat, nd, A)
    rng = np.random.default_rng(3)
    g = torch.tensor(rng.standard_normal((nd, nlat, 2 * nlat)), device="cuda")
    s = mesh.spectral_zeros()
    s_eager = mesh.spectral_zeros()
    mesh.reserve(nd)
    graph = jgcm.Step_Graph(mesh, lambda: jgcm.Trans_Grid_To_Spherical(mesh, g, s))
    # a call inside a capture that needs more scratch fails cleanly
    big = mesh.grid_zeros(8 * nd)
    bigs = mesh.spectral_zeros(8 * nd)
    with pytest.raises(RuntimeError, match="reserve it before capturing"):
        jgcm.Step_Graph(mesh, lambda: jgcm.Trans_Grid_To_Spherical(mesh, big, bigs, batch=True))
    # growth outside a capture while the graph lives: the graph keeps its buffers
    jgcm.Trans_Grid_To_Spherical(mesh, big, bigs, batch=True)
    jgcm.Trans_Grid_To_Spherical(mesh, g, s_eager)
    for _ in range(3):
        s.zero_()
        graph.launch()
        mesh.synchronize()
        assert torch.equal(s, s_eager)
    graph.close()
