// Serving loop interface (continuous batching over resident boards).
#pragma once
