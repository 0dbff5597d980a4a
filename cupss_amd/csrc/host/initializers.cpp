// initializers.cpp -- host-side initial conditions written into real_array
// (reference: src/initializers.cpp:14-267).  The formulas and the use of
// rand()/srand(time(0)) follow the reference, including its quirks:
// initializeUniformNoise draws value*0.01*(rand()%200-100);
// initializeNormalNoise uses sigma^2 as the standard deviation and fills
// pairs (index, index+1) from Box-Muller on rand()%100+1 uniforms.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>

#include "../../../include/cupss.h"

namespace {

void require_field(evolver *e, const std::string &name, const char *what)
{
    if (e->existsField(name) < 0) {
        std::cout << "ERROR in " << what << ", " << name << " not found" << std::endl;
        std::exit(1);
    }
}

}  // namespace

void evolver::initializeUniform(std::string f, float value)
{
    require_field(this, f, "initialize uniform");
    float2 *a = fieldsMap[f]->real_array;
    for (long idx = 0; idx < (long)sx * sy * sz; idx++) a[idx].x = value;
}

void evolver::initializeUniformNoise(std::string f, float value)
{
    require_field(this, f, "initialize uniform");
    srand(time(0));
    float2 *a = fieldsMap[f]->real_array;
    for (long idx = 0; idx < (long)sx * sy * sz; idx++) a[idx].x = value * 0.01f * (float)(rand() % 200 - 100);
}

void evolver::initializeNormalNoise(std::string f, float mean, float sigma)
{
    require_field(this, f, "initialize uniform");
    srand(time(0));
    float2 *a = fieldsMap[f]->real_array;
    const long n = (long)sx * sy * sz;
    for (long idx = 0; idx + 1 < n + 1; idx += 2) {
        const float v1 = 0.01 * (float)(rand() % 100 + 1);
        const float v2 = 0.01 * (float)(rand() % 100 + 1);
        const double r = std::sqrt(-2.0 * std::log(v1));
        a[idx].x = sigma * sigma * r * std::cos(2.0 * PI * v2) + mean;
        if (idx + 1 < n) a[idx + 1].x = sigma * sigma * r * std::sin(2.0 * PI * v2) + mean;
    }
}

void evolver::initializeHalfSystem(std::string f, float val1, float val2, float xi, int direction)
{
    require_field(this, f, "initialize half system");
    if (xi <= 0.0) {
        std::cout << "ERROR in initialize, interface width cannot be 0 or negative" << std::endl;
        std::exit(1);
    }
    if (direction < 1 || direction > 3) {
        std::cout << "ERROR in initialize, direction can be 1, 2 or 3 for x, y, z, respectively" << std::endl;
        std::exit(1);
    }
    srand(time(0));
    const int extent = direction == 1 ? sx : (direction == 2 ? sy : sz);
    float2 *a = fieldsMap[f]->real_array;
    for (int k = 0; k < sz; k++)
        for (int j = 0; j < sy; j++)
            for (int i = 0; i < sx; i++) {
                const int ref = direction == 1 ? i : (direction == 2 ? j : k);
                a[(long)k * sx * sy + (long)j * sx + i].x =
                    val1 + (val2 - val1) * 0.5 * (1.0 + std::tanh((ref - extent / 2) / (std::sqrt(2) * xi)));
            }
}

void evolver::initializeDroplet(std::string f, float val1, float val2, float radius, float xi, int p_x, int p_y,
                                int p_z)
{
    require_field(this, f, "initialize droplet");
    if (xi <= 0.0 || radius <= 0.0) {
        std::cout << "ERROR in initialize, droplet radius and interface width cannot be 0 or negative" << std::endl;
        std::exit(1);
    }
    const int cx = p_x % sx, cy = p_y % sy, cz = p_z % sz;
    srand(time(0));
    float2 *a = fieldsMap[f]->real_array;
    for (int k = 0; k < sz; k++)
        for (int j = 0; j < sy; j++)
            for (int i = 0; i < sx; i++) {
                const float xd = i - cx, yd = j - cy, zd = k - cz;
                const float r = std::sqrt(xd * xd + yd * yd + zd * zd);
                a[(long)k * sx * sy + (long)j * sx + i].x =
                    val1 + (val2 - val1) * 0.5 * (1.0 + std::tanh((r - radius) / (std::sqrt(2) * xi)));
            }
}

void evolver::addDroplet(std::string f, float val, float radius, float xi, int p_x, int p_y, int p_z)
{
    require_field(this, f, "initialize droplet");
    if (xi <= 0.0 || radius <= 0.0) {
        std::cout << "ERROR in initialize, droplet radius and interface width cannot be 0 or negative" << std::endl;
        std::exit(1);
    }
    const int cx = p_x % sx, cy = p_y % sy, cz = p_z % sz;
    srand(time(0));
    float2 *a = fieldsMap[f]->real_array;
    for (int k = 0; k < sz; k++)
        for (int j = 0; j < sy; j++)
            for (int i = 0; i < sx; i++) {
                const float xd = i - cx, yd = j - cy, zd = k - cz;
                const float r = std::sqrt(xd * xd + yd * yd + zd * zd);
                a[(long)k * sx * sy + (long)j * sx + i].x +=
                    val * 0.5 * (1.0 + std::tanh((radius - r) / (std::sqrt(2) * xi)));
            }
}

// CSV rows "x[, y[, z]], value" after `skiprows` header lines
// (src/initializers.cpp:216-267).
void evolver::initializeFromFile(std::string f, std::string file, int skiprows, char delimiter)
{
    require_field(this, f, "initialize droplet");
    FILE *chk = fopen(file.c_str(), "r");
    if (chk == NULL) {
        std::cout << "Initial file: " << file << " not found. Doing nothing" << std::endl;
        std::exit(1);
    }
    fclose(chk);
    std::ifstream in(file.c_str());
    std::string posx, posy = "0", posz = "0", value;
    for (int i = 0; i < skiprows; i++) std::getline(in, posx);
    float2 *a = fieldsMap[f]->real_array;
    const long n = (long)sx * sy * sz;
    for (long line = 0; line < n; line++) {
        if (!std::getline(in, posx, delimiter)) {
            std::cout << "Incompatible number of lines on file " << file << ", only " << line + 1
                      << " found for a system size of " << n << std::endl;
            std::exit(1);
        }
        if (dimension > 1) std::getline(in, posy, delimiter);
        if (dimension > 2) std::getline(in, posz, delimiter);
        std::getline(in, value);
        const long idx = (long)std::stoi(posz) * sx * sy + (long)std::stoi(posy) * sx + std::stoi(posx);
        a[idx].x = std::stof(value);
    }
}
