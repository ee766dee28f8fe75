"""Runtime HIP code generation for weak forms (the HIP backend of the element code
generator): expression IR (``expr``), scalar weak forms and their element kernels
(``scalar``), forms of the compiled kernels and others (``library``), and compilation with
hiprtc into the running library (``runtime``)."""
