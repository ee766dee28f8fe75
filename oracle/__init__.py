"""CPU oracle for the pyoomph assembly-and-solve path.

THIS PACKAGE IS TEST INFRASTRUCTURE. Only ``tests/``, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the checker (or as the
timed CPU baseline). The product package ``pyoomph_amd`` never imports it.

It restates, in plain Python / numpy / C, the reference algorithms on the hot path:

* ``meshes``     -- ``RectangularQuadMesh`` / ``CuboidBrickMesh`` templates, the C1->C2
                    template conversion and oomph's node / equation numbering
                    (pyoomph/meshes/simplemeshes.py:225-300, 456-545;
                    src/meshtemplate.cpp:401-408, 494-500, 579-621, 1378-1418;
                    src/mesh2d.hpp:265-300; oomph mesh.cc:686-708, nodes.cc:896-927).
* ``tables``     -- quadrature tables and Lagrange shape functions
                    (oomph integral.cc:50-53, 84-102, 169-206, 344-369;
                    Telements.h:523-542, 627-662; shape.h:607-650; Qelements.cc:320-372,
                    621-658; src/elements.cpp:9288-9305).
* ``oracle.c``   -- the generated element residual / Jacobian / dR/dparam functions of the
                    four weak forms, restated as plain C loops (src/codegen.cpp:3912-4109,
                    2015-2257; src/jitbridge_nohang.h:23-50; src/elements.cpp:3593-3990,
                    4564-4575), and oomph's ``vectors_of_pairs`` CSR builder
                    (oomph problem.cc:5459-5659).
* ``assembly``   -- element loop, residual assembly and CSR (oomph problem.cc:3810-3900,
                    4038-4170).
* ``faces``      -- boundary-face (interface) elements carrying NeumannBC(u=alpha*u + beta)
                    on P1 / P2 triangles (pyoomph/meshes/bcs.py:71-89; src/mesh.cpp:615-667;
                    oomph Telements.h TElementShape<1,2>, <1,3>; integral.cc:303-316).
* ``newton``     -- oomph's Newton loop with the SuperLU solve called the way pyoomph calls
                    it (oomph problem.cc:8986-9345; double_vector.cc:611-640;
                    pyoomph/solvers/scipy.py:47-104).

Parity status of each function is stated in its docstring ("pinned by ..." or
"parity unpinned").
"""
