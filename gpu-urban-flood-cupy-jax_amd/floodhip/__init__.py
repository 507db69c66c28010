"""floodhip: MI355X-native per-timestep solver for the unstructured urban flood model."""
