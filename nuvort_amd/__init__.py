"""MI355X-native filament velocity and timestep path of NuVort."""
